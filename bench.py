"""Benchmark: photons/s traced (source + dither + HRMA + grating + detector)
on one MI355X, with the CPU oracle timed on the host cores beside it.

One step = one device-resident chunk of a workload, created on the GPU and
traced through mirror, grating and detector without leaving HBM.  The
default workload (--config hetg-acis-s) is 1e7 photons of an on-axis point
source with a photon-index-2 power law over 0.4-8 keV, INTERNAL dither
(16 arcsec, 1000 / 707 s), HETG and ACIS-S, 50 ks of arrivals;
--config letg-hrc-s is 5e7 photons of a 2 arcmin disk, flat 0.07-10 keV,
LETG and HRC-S.  `value` is photons created and traced per second of wall
time over the timed steps.  --config batched-acis-i traces 4096 point
sources in one HIP-graph-launched pass; --config diff times gradient
evaluations of the differentiable mirror (gradient evals/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--photons N]
"""
import argparse
import json

import numpy as np
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Algorithmic FP64 work per photon entering each kernel (adds/subs, muls,
# divs and sqrts of the restated reference algorithm, counted by the
# op-counting oracle build on each workload:
# `python -m oracle.opcount.opcount --photons 100000 --workload <name>`).
# "grating" and "detector" are k_grating and k_detect.  Transcendental calls
# per photon are listed beside.
FLOPS_PER_PHOTON = {
    "hetg-acis-s": {"source": 146.0, "k_hrma_enter": 130.3, "k_hrma_reflect": 598.6,
                    "grating": 249.3, "detector": 200.1},
    "letg-hrc-s": {"source": 239.0, "k_hrma_enter": 130.3, "k_hrma_reflect": 442.0,
                   "grating": 1422.2, "detector": 278.0},
    "batched-acis-i": {"source": 11.0, "k_hrma_enter": 130.2, "k_hrma_reflect": 576.9,
                       "grating": 0.0, "detector": 172.4},
    # per photon of the pass (`--workload diff`): the weighted forward pass
    # with the tilt tangents (a source at infinity), split into its part up
    # to the paraboloid intersection (k_diff_enter, values only, every
    # photon) and the rest (k_diff_forward, the photons that get there), the
    # oracle's finite-difference reflectivity slope left out; and the adjoint
    "diff": {"k_diff_enter": 122.7, "k_diff_forward": 1804.2 - 122.7, "k_diff_adjoint": 42.7},
}
TRANSCENDENTALS_PER_PHOTON = {
    "hetg-acis-s": {"source": 11.0, "k_hrma_enter": 4.00, "k_hrma_reflect": 13.62,
                    "grating": 8.81, "detector": 0.69},
    "letg-hrc-s": {"source": 15.0, "k_hrma_enter": 4.00, "k_hrma_reflect": 10.09,
                   "grating": 13.69, "detector": 0.64},
    "batched-acis-i": {"source": 0.0, "k_hrma_enter": 4.00, "k_hrma_reflect": 13.11,
                       "grating": 0.0, "detector": 0.17},
    "diff": {"k_diff_enter": 4.00, "k_diff_forward": 11.91 - 4.00, "k_diff_adjoint": 0.0},
}
# Algorithmic HBM bytes per photon entering each kernel: the structure-of-
# arrays columns it must read and write, plus the active-list index in and
# the survivor index out.
BYTES_PER_PHOTON = {
    # energy, p, t, tag, dither, scan word (the position, shell and flags
    # are the mirror entry's to write)
    "source": 8 + 24 + 8 + 4 + 24 + 8,
    # in: index, tag, p (fresh photons: no flags read); out: x, p, shell,
    # flags, stream position, index
    "k_hrma_enter": (4 + 4 + 24) + (24 + 24 + 1 + 4 + 4 + 4),
    # in: index, flags, shell, tag, stream position, x, p, energy; out: x, p, flags, index
    "k_hrma_reflect": (4 + 4 + 1 + 4 + 4 + 24 + 24 + 8) + (24 + 24 + 4 + 4),
    "grating": (4 + 4 + 4 + 1 + 8 + 24 + 24 + 1) + (24 + 24 + 4 + 1 + 4),
    "detector": (4 + 4 + 4 + 8 + 8 + 24 + 24) + (24 + 24 + 4 + 1 + 4 + 4 + 4),
    # per photon entering: in: list index; out: w, y, z, their 3 x 5
    # derivatives, shell (the image stays in L2 / LDS)
    "k_diff_forward": 4 + 8 * (3 + 15) + 1,
    # per photon of weight != 0 (the others: their weight, 8 B): in: y, z,
    # the 15 derivatives, shell
    "k_diff_adjoint": 8 * (2 + 15) + 1,
}
# HBM bytes per photon entering a kernel, measured with rocprofv3 counters
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per dispatch, the factors that make a
# 1 GiB device copy read back 1 GiB each way); commands and figures in
# profiles/pmc_hetg_acis_s_1e7.md and profiles/pmc_letg_diff.md.  Absent
# when not measured for the current kernels.
# The figure is not measured inside this run (counters need their own
# profiler runs): each entry names the profile it comes from, and the bench
# line carries that name beside `traffic`.
TRAFFIC_BYTES_PER_PHOTON = {
    ("hetg-acis-s", "k_hrma_reflect"): (312.9, "profiles/pmc_hetg_acis_s_1e7.md"),
    ("letg-hrc-s", "k_hrma_reflect"): (321.2, "profiles/pmc_letg_diff.md"),
    ("letg-hrc-s", "grating"): (333.9, "profiles/pmc_letg_diff.md"),
    # per listed photon
    ("diff", "k_diff_forward"): (357.1, "profiles/pmc_letg_diff.md"),
}


def traffic_of(cfg, kernel, units):
    """(HBM bytes per launch, source profile) of a kernel, or (None, None)."""
    tb = TRAFFIC_BYTES_PER_PHOTON.get((cfg, kernel))
    if tb is None:
        return None, None
    return tb[0] * units, tb[1]
DESCRIPTION = {
    "hetg-acis-s": "HETG + ACIS-S: on-axis point source, power law index 2 over 0.4-8 keV (FILE table), "
                   "HEG/MEG orders from the CALDB efficiency tables, INTERNAL dither 16 arcsec "
                   "1000/707 s, 50 ks",
    "letg-hrc-s": "LETG + HRC-S: uniform disk of 2 arcmin radius, flat 0.07-10 keV, LETG orders from the "
                  "CALDB efficiency tables with fine and coarse support gratings, HRC-S with HESF, MCP "
                  "blur and pixel map, INTERNAL dither, 50 ks",
}
DESCRIPTION["batched-acis-i"] = (
    "Batched: 4096 point sources at off-axis angles uniform over 0-15 arcmin with random azimuths, "
    "photon counts log-uniform over 1e2-1e5 per source, power law index 2 over 0.3-10 keV, HRMA + "
    "ACIS-I, arrival times uniform over the exposure, all sources in one HIP-graph-launched pass "
    "with the autotuned schedule")
DESCRIPTION["diff"] = (
    "Differentiable: 20 misalignment parameters (tilt az/el, decenter y/z, despace x of the 4 shell "
    "pairs) drawn from N(0, 1 arcsec / 10 um), 1e7 survival-weighted photons of an on-axis point "
    "source at 1.49 keV through the mirror, loss = squared difference of FP32-accumulated 256x256 "
    "focal-plane images (0.004 mm pixels) against the image at perturbed parameters")
DESCRIPTION["output"] = (
    "Output files: the column files (marx_write_photons format, every column) of one traced HETG + "
    "ACIS-S chunk of 1e7 photons, written from HBM by mxg_write_device, against downloading the chunk "
    "and writing it with marx_write_photons")
DESCRIPTION["sources"] = (
    "Source models: 1e7 photons per step of an on-axis POINT, DISK (2 arcmin), GAUSS (60 arcsec), BETA "
    "(10 arcsec core, beta 0.7), LINE (30 arcmin) and IMAGE (512 x 512 pixels of 0.5 arcsec: a 20 arcsec "
    "Gaussian blob on a 1 percent floor) source, flat 0.3-10 keV, INTERNAL dither, created on the device "
    "and traced through HRMA + ACIS-S")
DESCRIPTION["events"] = (
    "Event coordinates: marx2fits's tiled, focal-plane and sky pixels, flight grades and EDSER "
    "subpixel offsets of the detected photons of one traced HETG + ACIS-S chunk of 1e7 photons "
    "(dither held per 3.241 s frame), computed in HBM by mxg_event_coords")
DEFAULT_PHOTONS = {"hetg-acis-s": 10_000_000, "letg-hrc-s": 50_000_000, "batched-acis-i": 0,
                   "diff": 10_000_000, "output": 10_000_000, "sources": 10_000_000, "events": 10_000_000,
                   "host-aos": 2_000_000}
# algorithmic HBM bytes per event row: k_event_heads reads the row's list
# entry (4) and time (8) and writes its frame head (4); k_event_coords reads
# the list entry and head (8), tag, chip id, chip x/y and PI (4+1+4+4+4; the
# held dither state is one read per frame) and writes nine columns (50)
EVENT_BYTES = {"k_event_heads": 16, "k_event_coords": 75}
SOURCE_SHAPES = {"POINT": {}, "DISK": {"S-DiskTheta0": 0.0, "S-DiskTheta1": 120.0},
                 "GAUSS": {"S-GaussSigma": 60.0}, "BETA": {"S-BetaCoreRadius": 10.0, "S-BetaBeta": 0.7},
                 "LINE": {"S-LineTheta": 1800.0, "S-LinePhi": 30.0}, "IMAGE": {}}
# the IMAGE shape's synthetic image: 512 x 512 pixels of 0.5 arcsec, a
# Gaussian blob of 20 arcsec sigma over a flat floor of 1 percent of its peak
IMAGE_SHAPE = (512, 512, 0.5)
DIFF_NPIX, DIFF_PIXEL_MM = 256, 0.004
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
# The path does no matrix work: its compute roof is the FP64 vector ALU.
VALU_PIPE = ("FP64 vector ALU (VALU; no matrix instructions on this path): 78.6 TFLOP/s is MI355X's "
             "FP64 vector peak")
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E
STAGES = ["source", "mirror", "grating", "detector"]


def cpu_baseline(tracer, target_seconds=20.0):
    """The FP64 oracle over a bounded sample of the same workload, OpenMP
    over the host cores (one thread per chunk of 50k photons).  Two probe
    runs size the timed sample to about target_seconds."""
    from oracle.oracle import Oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    o = Oracle(tracer)
    chunk = 50_000

    def timed(n):
        t0 = time.perf_counter()
        if o.trace_parallel(n, chunk, threads) < 0:
            raise RuntimeError("oracle failed")
        return time.perf_counter() - t0

    n = chunk * threads
    for target in (0.5, 2.0):
        dt = timed(n)
        n = int(min(max(n * target / max(dt, 1e-3), chunk * threads), 4e8))
    sample = int(min(max(n * target_seconds / 2.0, chunk * threads), 4e8))
    sample -= sample % chunk
    dt = timed(sample)
    return {"value": sample / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{sample} photons of the same workload (FP64 CPU oracle, chunks of {chunk}, "
                      f"{threads} OpenMP threads), {dt:.1f} s"}


def kernel_roofline(cfg, kernels):
    """The roofline object for the slowest of the timed kernels."""
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    units = kernels[dom]["photons_in"]
    traffic, traffic_src = traffic_of(cfg, dom, units)
    return {"bound": "valu", "pipe": VALU_PIPE,
            "kernel": "k_grating (stage)" if dom == "grating" else
                      "k_detect (stage)" if dom == "detector" else dom,
            "achieved": kernels[dom]["tflops"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": kernels[dom]["tflops"] / PEAK_FP64_TFLOPS,
            "traffic": traffic, "traffic_source": traffic_src,
            "flops_per_photon": FLOPS_PER_PHOTON[cfg][dom],
            "transcendentals_per_photon": TRANSCENDENTALS_PER_PHOTON[cfg][dom],
            "photons_per_launch": units, "launch_ms": kernels[dom]["ms"],
            "hbm": {"achieved": kernels[dom]["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": kernels[dom]["gbs"] / PEAK_HBM_GBS,
                    "bytes_per_photon": BYTES_PER_PHOTON[dom]}}


def measure_batched(args, tmp, torch):
    """Batched point sources: one step is one HIP-graph-launched pass over
    every source (mxg_trace_batch) with the schedule the autotuner picked;
    beside it the plain per-source GPU loop (one uncaptured pass per source)
    over the same photons, and the CPU oracle per source on the host cores."""
    from marx_amd.workloads import batched_acis_i
    cfg = "batched-acis-i"
    tracer, dirs, counts = batched_acis_i(tmp)
    N, T, K = int(counts.sum()), args.exposure, args.steps
    progress = lambda msg: print(f"bench batched: {msg}", file=sys.stderr, flush=True)
    progress(f"{len(counts)} sources, {N} photons; autotuning")
    sched, tuned_ms = tracer.autotune_batch(dirs, counts, exposure=T, repeats=2)
    progress(f"schedule {sched}, graph pass {tuned_ms:.2f} ms")

    def step(k):
        return tracer.trace_batch(dirs, counts, tag_start=(k * N) % (1 << 31), exposure=T, use_graph=True)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    detected = 0
    t0 = time.perf_counter()
    for k in range(K):
        detected += step(args.warmup + k)[0]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = K * N / elapsed
    progress(f"{K} graph-launched steps: {elapsed / K * 1e3:.2f} ms per step")

    # per-kernel device times from one uncaptured pass of the same batch
    # under the staged schedule (the persistent pass is one kernel), which
    # the per-source loop below also uses
    tracer.set_schedule(persistent_tile=0)
    tracer.trace_batch(dirs, counts, exposure=T, use_graph=False)
    sms = tracer.stage_ms()
    kms, kin = tracer.kernel_stats()
    c = tracer.stage_counts()
    kernels = {}
    for name, ms_k, n_k in (("k_hrma_enter", kms[0], kin[0]), ("k_hrma_reflect", kms[1], kin[1]),
                            ("detector", sms[3], c[1])):
        t = ms_k * 1e-3
        kernels[name] = {"ms": ms_k, "photons_in": n_k,
                         "tflops": FLOPS_PER_PHOTON[cfg][name] * n_k / t / 1e12 if t > 0 else None,
                         "gbs": BYTES_PER_PHOTON[name] * n_k / t / 1e9 if t > 0 else None}

    # the plain per-source GPU loop over the same photons
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(len(counts)):
        tracer.trace_batch(dirs[k:k + 1], counts[k:k + 1], tag_start=int(offs[k]), exposure=T, use_graph=False)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    progress(f"per-source loop: {loop_s:.2f} s")
    tracer.set_schedule(persistent_tile=sched["persistent_tile"],
                        persistent_blocks_per_cu=sched["persistent_blocks_per_cu"])

    out = {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": value, "unit": "photons/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": elapsed / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCRIPTION[cfg], "name": cfg, "sources": len(counts),
                   "photons_per_step": N, "exposure_s": T, "detected_per_step": detected / K},
        "schedule": sched, "autotuned_pass_ms": tuned_ms,
        "pass": "persistent" if sched["persistent_tile"] else "staged",
        "per_source_loop": {"value": N / loop_s, "unit": "photons/s", "seconds": loop_s,
                            "sources": len(counts)},
        "batched_over_loop": value / (N / loop_s),
        "roofline": kernel_roofline(cfg, kernels),
        "stages_ms": {"source": sms[0], "mirror": sms[1], "detector": sms[3]},
        "kernels": kernels,
    }
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline_batch(tracer, dirs, counts, T)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def measure_diff(args, torch):
    """Differentiable mirror: one step is one gradient evaluation through the
    torch.autograd.Function (weighted forward pass with tangents, FP32 loss,
    loss.backward() running the adjoint kernel).  Beside it: the plain
    alternative on the same GPU, a central-difference gradient from 40
    value-only forward passes (mxg_diff_image), with its error against the
    adjoint gradient; and the FP64 CPU oracle (forward and adjoint, OpenMP)
    on a sample of the photons."""
    from marx_amd.diff import NUM_PARAMS, DiffMirror
    from marx_amd.workloads import diff_mirror
    cfg, K = "diff", args.steps
    N = args.photons or DEFAULT_PHOTONS[cfg]
    progress = lambda msg: print(f"bench diff: {msg}", file=sys.stderr, flush=True)
    tracer, p0, pt = diff_mirror()
    dm = DiffMirror(tracer, photons=N, tag_start=0, npix=DIFF_NPIX, pixel_size=DIFF_PIXEL_MM)
    with torch.no_grad():
        target = dm.image(torch.as_tensor(pt, dtype=torch.float64)).clone()

    def step():
        params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
        loss = ((dm.image(params) - target) ** 2).sum()
        loss.backward()
        return params.grad, loss.detach()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    fwd_ms = adj_ms = ent_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(K):
        grad, loss = step()
        f, a, e, listed = dm.kernel_ms()
        fwd_ms += f
        adj_ms += a
        ent_ms += e
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = K / elapsed
    grad = grad.detach().numpy().copy()
    loss = float(loss)
    weighted = int((dm.state(tangents=False)["w"] != 0.0).sum())
    progress(f"{K} gradient evaluations: {elapsed / K * 1e3:.2f} ms each")

    # the plain loop: central differences of value-only passes
    img = torch.empty((DIFF_NPIX, DIFF_NPIX), dtype=torch.float32, device="cuda")
    t64 = target.double()

    def loss_at(p):
        dm.forward_into(p, img, tangents=False)
        return float(((img.double() - t64) ** 2).sum())

    steps_fd = np.where(np.arange(NUM_PARAMS) % 5 < 2, 1e-2, 1e-4)    # arcsec, mm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fd = np.zeros(NUM_PARAMS)
    for j in range(NUM_PARAMS):
        dp = np.zeros(NUM_PARAMS)
        dp[j] = steps_fd[j]
        fd[j] = (loss_at(p0 + dp) - loss_at(p0 - dp)) / (2 * steps_fd[j])
    torch.cuda.synchronize()
    fd_s = time.perf_counter() - t0
    value_ms, _, value_enter_ms, _ = dm.kernel_ms()
    progress(f"finite-difference gradient (40 value-only passes): {fd_s * 1e3:.1f} ms")

    # per-kernel figures: algorithmic FLOPs per photon of the pass x photons
    # of the pass; bytes per photon entering x photons entering
    kernels = {}
    byt = {"k_diff_enter": None,
           "k_diff_forward": BYTES_PER_PHOTON["k_diff_forward"] * listed,
           "k_diff_adjoint": 8 * N + BYTES_PER_PHOTON["k_diff_adjoint"] * weighted}
    for name, ms_k, n_in in (("k_diff_enter", ent_ms / K, N), ("k_diff_forward", fwd_ms / K, listed),
                             ("k_diff_adjoint", adj_ms / K, N)):
        t = ms_k * 1e-3
        kernels[name] = {"ms": ms_k, "photons_in": n_in,
                         "flops_per_pass_photon": FLOPS_PER_PHOTON[cfg][name],
                         "tflops": FLOPS_PER_PHOTON[cfg][name] * N / t / 1e12 if t > 0 else None,
                         "gbs": byt[name] / t / 1e9 if (t > 0 and byt[name]) else None}
    kernels["k_diff_enter"]["note"] = "with the list compaction after it"
    kernels["value_only_pass"] = {"ms": value_enter_ms + value_ms, "k_diff_forward<0>_ms": value_ms,
                                  "k_diff_enter_ms": value_enter_ms}
    fk = kernels["k_diff_forward"]
    traffic, traffic_src = traffic_of(cfg, "k_diff_forward", listed)
    rf = {"bound": "valu", "pipe": VALU_PIPE,
          "kernel": "k_diff_forward<2>", "achieved": fk["tflops"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
          "frac": fk["tflops"] / PEAK_FP64_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
          "flops_per_pass_photon": FLOPS_PER_PHOTON[cfg]["k_diff_forward"],
          "transcendentals_per_pass_photon": TRANSCENDENTALS_PER_PHOTON[cfg]["k_diff_forward"],
          "pass_photons": N, "photons_per_launch": listed, "launch_ms": fk["ms"],
          "hbm": {"achieved": fk["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": fk["gbs"] / PEAK_HBM_GBS,
                  "bytes_per_photon": BYTES_PER_PHOTON["k_diff_forward"]}}
    adj = kernels["k_diff_adjoint"]
    out = {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": value, "unit": "gradient evals/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": elapsed / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCRIPTION[cfg], "name": cfg, "photons_per_eval": N, "params": NUM_PARAMS,
                   "image": f"{DIFF_NPIX}x{DIFF_NPIX} FP32, {DIFF_PIXEL_MM} mm pixels",
                   "total_weight": dm.total_weight, "loss": loss},
        "photons_per_s": value * N,
        "roofline": rf,
        "adjoint_hbm": {"achieved": adj["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": adj["gbs"] / PEAK_HBM_GBS if adj["gbs"] else None,
                        "bytes": byt["k_diff_adjoint"], "weighted_photons": weighted},
        "kernels": kernels,
        "fd_loop": {"value": 1.0 / fd_s, "unit": "gradient evals/s", "passes": 2 * NUM_PARAMS,
                    "steps": "1e-2 arcsec, 1e-4 mm", "seconds": fd_s,
                    "max_abs_err_over_max_grad": float(np.abs(fd - grad).max() / np.abs(grad).max())},
        "gradient_over_fd_loop": value * fd_s,
    }
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline_diff(tracer, p0, target.cpu().numpy(), N)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def cpu_baseline_diff(tracer, params, target, photons, target_seconds=20.0):
    """The FP64 oracle's forward and adjoint passes (OpenMP over the host
    cores) over a sample of the photons, scaled to gradient evaluations of
    `photons` photons per second.  Two probes size the sample to about
    target_seconds: a part of the photons, or whole evaluations repeated."""
    from marx_amd._lib import ImageSpec
    from marx_amd.diff import setup
    from oracle.oracle import Oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    o = Oracle(tracer)
    spec = ImageSpec(DIFF_NPIX, DIFF_PIXEL_MM, 0.0, 0.0)
    d = setup(params)
    t64 = np.asarray(target, dtype=np.float64)

    def timed(n):
        t0 = time.perf_counter()
        img, st = o.diff_forward(d, n, 0, spec, nthreads=threads)
        o.diff_adjoint(st, spec, 2.0 * (img - t64), nthreads=threads)
        return time.perf_counter() - t0

    n = 20_000 * threads
    for target_s in (0.5, 2.0):
        dt = timed(n)
        rate = n / max(dt, 1e-3)
        n = int(min(max(rate * target_s, 1000 * threads), photons))
    if rate * target_seconds < photons:
        sample = int(max(rate * target_seconds, 1000 * threads))
        dt = timed(sample)
        what = f"the first {sample} of the {photons} photons"
        evals = sample / photons
    else:
        reps = int(max(1, min(5, rate * target_seconds // photons)))
        t0 = time.perf_counter()
        for _ in range(reps):
            timed(photons)
        dt = time.perf_counter() - t0
        what = f"{reps} whole evaluation(s) of the {photons} photons"
        evals = reps
    return {"value": evals / dt, "unit": "gradient evals/s", "cores": threads, "kind": "port",
            "sample": f"{what} (FP64 CPU oracle forward with tangents and adjoint, {threads} OpenMP "
                      f"threads), {dt:.1f} s"}


def measure_output(args, tmp, torch):
    """Output files of one traced chunk: one step writes every column file
    of the chunk's detected photons from HBM (mxg_write_device: rows
    selected and bytes packed on the GPU, one copy to the host, one write
    per file).  Beside it the AoS route on the same chunk: download the list
    and write it with marx_write_photons.  Files go to a temporary directory
    (page cache; the figures are not disk speeds)."""
    from marx_amd import Photons
    from marx_amd._lib import WRITE_ALL
    from marx_amd.workloads import hetg_acis_s
    cfg, K = "output", args.steps
    N = args.photons or DEFAULT_PHOTONS[cfg]
    progress = lambda msg: print(f"bench output: {msg}", file=sys.stderr, flush=True)
    t = hetg_acis_s(tmp, photons=N, exposure=args.exposure)
    ncol, ndet, _ = t.trace_device(N)
    a, b = os.path.join(tmp, "device"), os.path.join(tmp, "aos")
    os.makedirs(a)
    os.makedirs(b)
    for _ in range(args.warmup):
        t.write_device(a)
    t0 = time.perf_counter()
    for _ in range(K):
        rows = t.write_device(a)
    dev_s = (time.perf_counter() - t0) / K
    nbytes = sum(os.path.getsize(os.path.join(a, f)) for f in os.listdir(a))
    progress(f"{rows} rows, {nbytes / 1e6:.1f} MB per write; device path {dev_s * 1e3:.1f} ms")
    p = Photons(N)
    reps = max(1, min(K, 3))
    t0 = time.perf_counter()
    for _ in range(reps):
        t.download(p)
        p.write(b)
    aos_s = (time.perf_counter() - t0) / reps
    progress(f"download + marx_write_photons {aos_s * 1e3:.1f} ms")
    same = all(open(os.path.join(a, f), "rb").read() == open(os.path.join(b, f), "rb").read()
               for f in os.listdir(a))
    return {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": rows / dev_s, "unit": "rows/s written (all columns)", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": dev_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCRIPTION[cfg], "name": cfg, "photons_traced": ncol, "rows": rows,
                   "files": len(os.listdir(a)), "bytes_per_write": nbytes},
        "file_gb_per_s": nbytes / dev_s / 1e9,
        "aos_route": {"ms": aos_s * 1e3, "value": rows / aos_s, "unit": "rows/s written"},
        "device_over_aos_route": aos_s / dev_s,
        "files_identical": same,
    }


def measure_host_aos(args, tmp, torch):
    """The libmarx drop-in route on one chunk, as marx.c drives it: each step
    is marx_create_photons into host AoS records, then marx_mirror_reflect
    (upload), marx_grating_diffract and marx_detect (download into the
    caller's records).  Wall time per chunk, PCIe copies and the AoS <-> SoA
    conversions included; the device-resident rate is the default line's."""
    from marx_amd import Photons
    from marx_amd.workloads import hetg_acis_s
    cfg, K = "host-aos", args.steps
    N = args.photons or DEFAULT_PHOTONS[cfg]
    progress = lambda msg: print(f"bench host-aos: {msg}", file=sys.stderr, flush=True)
    t = hetg_acis_s(tmp, photons=N, exposure=args.exposure)
    ph = Photons(N)

    calls = {"create": 0.0, "reflect": 0.0, "diffract": 0.0, "detect": 0.0}

    def step(timed=False):
        # every call returns with the caller's records valid (each one
        # synchronises), so host clocks between them time each call
        c0 = time.perf_counter()
        made = t.create_photons(ph, N)
        c1 = time.perf_counter()
        t.reflect(ph)
        c2 = time.perf_counter()
        t.diffract(ph)
        c3 = time.perf_counter()
        t.detect(ph)
        c4 = time.perf_counter()
        if timed:
            for k, d in zip(calls, (c1 - c0, c2 - c1, c3 - c2, c4 - c3)):
                calls[k] += d
        return made

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    made = 0
    for _ in range(K):
        made += step(timed=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / K
    made /= K
    calls_ms = {k: v / K * 1e3 for k, v in calls.items()}
    progress(f"{made:.0f} photons per chunk, {dt * 1e3:.1f} ms per chunk, calls {calls_ms}")
    return {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": made / dt, "unit": "photons/s (host AoS in and out, PCIe included)", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCRIPTION["hetg-acis-s"], "name": cfg, "photons_per_step": made,
                   "route": "marx_create_photons + marx_mirror_reflect + marx_grating_diffract + marx_detect"},
        "calls_ms": calls_ms,
    }


def measure_events(args, tmp, torch):
    """Event coordinates of one traced chunk: each step computes every event
    column of the chunk's detected photons (mxg_event_coords, EDSER) in HBM;
    the device time of its kernels (frame heads and coordinates) is the
    figure, the call's wall time with the copies of the columns to host
    memory is reported beside it.  CPU baseline: the oracle's row loop (one
    thread) over the first rows of the same columns, as written to the
    dump files."""
    from marx_amd.workloads import hetg_acis_s
    cfg, K = "events", args.steps
    N = args.photons or DEFAULT_PHOTONS[cfg]
    progress = lambda msg: print(f"bench events: {msg}", file=sys.stderr, flush=True)
    t = hetg_acis_s(tmp, photons=N, exposure=args.exposure)
    ncol, ndet, _ = t.trace_device(N)
    for _ in range(args.warmup):
        t.event_coords(ndet)
    kern = 0.0
    t0 = time.perf_counter()
    for _ in range(K):
        out = t.event_coords(ndet)
        kern += t.event_kernel_ms()
    wall = (time.perf_counter() - t0) / K
    kern /= K
    rows = len(out["tag"])
    progress(f"{rows} rows; kernels {kern:.3f} ms, call with host copies {wall * 1e3:.1f} ms")
    res = {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": rows / (kern * 1e-3), "unit": "event rows/s (device)", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": kern, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCRIPTION[cfg], "name": cfg, "photons_traced": ncol, "rows": rows,
                   "pix_adj": "EDSER"},
        "call_ms_with_host_copies": wall * 1e3,
        "roofline": {"bound": "hbm", "achieved": rows * sum(EVENT_BYTES.values()) / (kern * 1e-3) / 1e9,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "traffic": None,
                     "kernels": "k_event_heads + k_event_chunk_prefix + k_event_coords"},
    }
    res["roofline"]["frac"] = res["roofline"]["achieved"] / PEAK_HBM_GBS
    if not args.no_cpu:
        from oracle import dumpfile as D
        from oracle.oracle import Oracle
        from oracle.subpix import subpix_tables
        from marx_amd._lib import DATA_DIR
        d = os.path.join(tmp, "cols")
        os.makedirs(d)
        t.write_device(d)
        m = min(rows, 2_000_000)
        col = lambda f: D.read_column(open(os.path.join(d, f), "rb").read())[4][:m]
        cols = {"TIME": col("time.dat"), "TAG": col("tag.dat"), "CCDID": col("detector.dat"),
                "CHIPX": col("xpixel.dat"), "CHIPY": col("ypixel.dat"), "B_ENERGY": col("b_energy.dat")}
        for k, f in (("RA", "sky_ra.dat"), ("DEC", "sky_dec.dat"), ("ROLL", "sky_roll.dat"),
                     ("DET_DY", "det_dy.dat"), ("DET_DZ", "det_dz.dat"), ("DET_THETA", "det_theta.dat")):
            cols[k] = col(f)
        tables = subpix_tables(os.path.join(DATA_DIR, "caldb", "acisD1999-07-22subpixN0001_marx.fits"))
        o = Oracle(t)
        ec = t.event_cfg("EDSER")
        t0 = time.perf_counter()
        e = o.event_coords(ec, cols, *tables)
        cpu_s = time.perf_counter() - t0
        same = all(np.array_equal(e[k], out[k][:m]) for k in ("tag", "tdetx", "tdety", "fltgrade"))
        res["cpu_baseline"] = {"value": m / cpu_s, "unit": "event rows/s", "cores": 1, "kind": "port",
                               "sample": f"first {m} of the {rows} rows, oracle row loop (marx2fits order)",
                               "integer_columns_identical": bool(same)}
        progress(f"cpu baseline {m / cpu_s:.3g} rows/s")
    return res


def _write_source_image(path):
    ny, nx, pix = IMAGE_SHAPE
    yy, xx = np.mgrid[0:ny, 0:nx]
    r2 = ((xx - nx / 2) ** 2 + (yy - ny / 2) ** 2) * (pix / 20.0) ** 2
    img = (np.exp(-0.5 * r2) + 0.01).astype(">f4")
    cards = [("SIMPLE", "T"), ("BITPIX", "-32"), ("NAXIS", "2"), ("NAXIS1", str(nx)), ("NAXIS2", str(ny)),
             ("CDELT1", repr(-pix / 3600)), ("CDELT2", repr(pix / 3600))]
    hdr = "".join(f"{k:<8}= {v:>20}".ljust(80) for k, v in cards) + "END".ljust(80)
    hdr = hdr.ljust((len(hdr) + 2879) // 2880 * 2880)
    body = img.tobytes()
    with open(path, "wb") as f:
        f.write(hdr.encode("ascii") + body + b"\0" * ((-len(body)) % 2880))
    return path


def measure_sources(args, torch):
    """Source models: per shape, steps of 1e7 photons created on the device
    (k_source, k_scan_tiles, k_source_finish) and traced through HRMA and
    ACIS-S; the source stage's device time and the whole step's."""
    from marx_amd import Tracer
    from marx_amd.workloads import ON_AXIS
    cfg, K = "sources", args.steps
    N = args.photons or DEFAULT_PHOTONS[cfg]
    res = {}
    image = _write_source_image(os.path.join(tempfile.mkdtemp(prefix="marx_bench_"), "image.fits"))
    for shape, par in SOURCE_SHAPES.items():
        if shape == "IMAGE":
            par = {"S-ImageFile": image}
        t = Tracer(seed=1, GratingType="NONE", DetectorType="ACIS-S", SourceType=shape, SpectrumType="FLAT",
                   MinEnergy=0.3, MaxEnergy=10.0, DitherModel="INTERNAL", **par, **ON_AXIS)
        t.init_source(SourceFlux=N / (args.exposure * t.geometric_area))
        for k in range(args.warmup):
            t.trace_device(N, tag_start=k * N, start_time=k * args.exposure)
        torch.cuda.synchronize()
        src_ms = 0.0
        t0 = time.perf_counter()
        for k in range(K):
            _, _, ms = t.trace_device(N, tag_start=k * N, start_time=k * args.exposure, timing=True)
            src_ms += ms[0]
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / K
        res[shape] = {"step_ms": el * 1e3, "photons_per_s": N / el, "source_stage_ms": src_ms / K,
                      "source_photons_per_s": N / (src_ms / K * 1e-3)}
        print(f"bench sources: {shape} {res[shape]}", file=sys.stderr, flush=True)
    v = min(r["photons_per_s"] for r in res.values())
    return {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": v, "unit": "photons/s (slowest shape)", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": max(r["step_ms"] for r in res.values()), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": DESCRIPTION[cfg], "name": cfg, "photons_per_step": N},
        "shapes": res,
    }


def cpu_baseline_batch(tracer, dirs, counts, exposure, target_seconds=20.0):
    """The FP64 oracle source by source (one OpenMP task per source) over the
    first sources of the batch, as many as take about target_seconds."""
    from oracle.oracle import Oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    o = Oracle(tracer)

    def timed(m):
        t0 = time.perf_counter()
        if o.trace_batch_parallel(dirs[:m], counts[:m], 0, exposure, threads) < 0:
            raise RuntimeError("oracle failed")
        return time.perf_counter() - t0

    cum = np.cumsum(counts.astype(np.int64))

    def sources_for(photons):
        return int(min(len(counts), max(threads, np.searchsorted(cum, photons) + 1)))

    m = sources_for(cum[-1] * 0.001)
    for target in (0.5, 2.0):
        dt = timed(m)
        rate = cum[m - 1] / max(dt, 1e-3)
        m = sources_for(rate * target)
    m = sources_for(rate * target_seconds)
    dt = timed(m)
    n = int(counts[:m].sum())
    return {"value": n / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"the first {m} sources of the batch, {n} photons (FP64 CPU oracle, one OpenMP "
                      f"task per source, {threads} threads), {dt:.1f} s"}


GRID_CANDIDATES = {"enter": (4, 5, 8, 10, 15, 20), "reflect": (4, 8, 12, 16),
                   "grating": (0, 2, 3, 4, 6, 8, 12), "detect": (1, 2, 4, 8)}


def sweep_grids(args, tracer, step):
    """Each stage kernel's time over candidate blocks per CU, one kernel at a
    time with the others at the library's defaults (grating 0: one round of
    the blocks a CU holds at once).  The mirror's kernels are timed alone
    with HIP events, the grating and detector as their stages."""
    default = tracer.get_schedule()
    k = 0
    rows = []
    try:
        for kern, values in GRID_CANDIDATES.items():
            for v in values:
                tracer.set_schedule(**{kern: v})
                for _ in range(args.warmup):
                    step(k)
                    k += 1
                ms = 0.0
                for _ in range(args.steps):
                    _, _, stage = step(k, timing=True)
                    k += 1
                    kms, _ = tracer.kernel_stats()
                    ms += {"enter": kms[0], "reflect": kms[1], "grating": stage[3], "detect": stage[4]}[kern]
                rows.append({"kernel": kern, "blocks_per_cu": v, "ms": round(ms / args.steps, 4)})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
                tracer.set_schedule(**{kern: default[kern]})
    finally:
        tracer.set_schedule(**{kk: default[kk] for kk in ("source", "enter", "reflect", "grating", "detect")})
    return {"sweep": "blocks per CU of each stage kernel", "config": args.config, "steps": args.steps,
            "warmup": args.warmup, "default": default, "rows": rows}


def run_quietly(fn):
    """Runs fn with file descriptor 1 sent to stderr, so that the library's
    progress messages do not mix with the one JSON line on stdout."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn()
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def main():
    print(json.dumps(run_quietly(measure)))


def measure():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="hetg-acis-s", choices=sorted(DEFAULT_PHOTONS))
    ap.add_argument("--photons", type=int, default=0, help="photons per step (default: the config's)")
    ap.add_argument("--exposure", type=float, default=50_000.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--sweep-grids", action="store_true",
                    help="instead of the bench line: each stage kernel's HIP-event time over candidate "
                         "blocks per CU (the others at their defaults), one JSON line per candidate")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the steps, copy a MIB-MiB device buffer three times (the known-size "
                         "reference that HBM-traffic counters are scaled by)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): photons are "
              "independent, a multi-GPU run would only shard photon ids and add one final "
              "reduction, so the benchmark measures one GPU's share.", file=sys.stderr)
        sys.exit(2)

    import torch
    import marx_amd
    from marx_amd.workloads import WORKLOADS

    lib = marx_amd.load()
    if lib.mxg_init() != 0:
        raise RuntimeError("GPU init failed: " + marx_amd._lib.last_error())

    tmp = tempfile.mkdtemp(prefix="marx_bench_")
    if args.config == "batched-acis-i":
        return measure_batched(args, tmp, torch)
    if args.config == "diff":
        return measure_diff(args, torch)
    if args.config == "output":
        return measure_output(args, tmp, torch)
    if args.config == "sources":
        return measure_sources(args, torch)
    if args.config == "events":
        return measure_events(args, tmp, torch)
    if args.config == "host-aos":
        return measure_host_aos(args, tmp, torch)
    cfg = args.config
    N, T = (args.photons or DEFAULT_PHOTONS[cfg]), args.exposure
    tracer = WORKLOADS[cfg](tmp, photons=N, exposure=T)

    def step(k, timing=False):
        return tracer.trace_device(N, tag_start=(k * N) % (1 << 31), start_time=k * T, timing=timing)

    if args.sweep_grids:
        return sweep_grids(args, tracer, step)
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    stage_ms = {s: 0.0 for s in STAGES}
    stage_in = {s: 0 for s in STAGES}
    kern_ms = {"k_hrma_enter": 0.0, "k_hrma_reflect": 0.0}
    kern_in = {"k_hrma_enter": 0, "k_hrma_reflect": 0}
    detected = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        ncol, ndet, ms = step(args.warmup + k, timing=True)
        c = tracer.stage_counts()
        kms, kin = tracer.kernel_stats()
        for j, name in enumerate(("k_hrma_enter", "k_hrma_reflect")):
            kern_ms[name] += kms[j]
            kern_in[name] += kin[j]
        stage_ms["source"] += ms[0]
        stage_ms["mirror"] += ms[2]
        stage_ms["grating"] += ms[3]
        stage_ms["detector"] += ms[4]
        stage_in["source"] += N
        stage_in["mirror"] += c[0]
        stage_in["grating"] += c[1]
        stage_in["detector"] += c[2]
        detected += ndet
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    K = args.steps
    value = K * N / elapsed

    # per-kernel figures; the mirror's kernels are timed alone with HIP events
    # on the library's stream, grating and detector are their stage (one
    # kernel plus the list compaction)
    kernels = {}
    for name, ms_k, n_k in (("k_hrma_enter", kern_ms["k_hrma_enter"], kern_in["k_hrma_enter"]),
                            ("k_hrma_reflect", kern_ms["k_hrma_reflect"], kern_in["k_hrma_reflect"]),
                            ("grating", stage_ms["grating"], stage_in["grating"]),
                            ("detector", stage_ms["detector"], stage_in["detector"])):
        t = ms_k / K * 1e-3
        kernels[name] = {"ms": ms_k / K, "photons_in": n_k / K,
                         "tflops": FLOPS_PER_PHOTON[cfg][name] * n_k / K / t / 1e12 if t > 0 else None,
                         "gbs": BYTES_PER_PHOTON[name] * n_k / K / t / 1e9 if t > 0 else None}

    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    t_launch = kernels[dom]["ms"] * 1e-3
    units = kernels[dom]["photons_in"]
    achieved = kernels[dom]["tflops"]
    achieved_bw = kernels[dom]["gbs"]
    traffic, traffic_src = traffic_of(cfg, dom, units)

    out = {
        "metric": "photons/s traced (HRMA+grating+detector) 1 GPU vs host CPU; gradient evals/s",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": DESCRIPTION[cfg], "name": cfg,
                   "photons_per_step": N, "exposure_s": T,
                   "detected_per_step": detected / K},
        "roofline": {"bound": "valu", "pipe": VALU_PIPE,
                     "kernel": "k_grating (stage)" if dom == "grating" else
                               "k_detect (stage)" if dom == "detector" else dom,
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "flops_per_photon": FLOPS_PER_PHOTON[cfg][dom],
                     "transcendentals_per_photon": TRANSCENDENTALS_PER_PHOTON[cfg][dom],
                     "photons_per_launch": units, "launch_ms": t_launch * 1e3,
                     "hbm": {"achieved": achieved_bw, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": achieved_bw / PEAK_HBM_GBS,
                             "bytes_per_photon": BYTES_PER_PHOTON[dom]}},
        "stages_ms": {s: stage_ms[s] / K for s in STAGES},
        "stage_photons_in": {s: stage_in[s] / K for s in STAGES},
        "kernels": kernels,
    }
    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = 2 * src.numel() * 4
    if not args.no_cpu:
        cb = cpu_baseline(tracer)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    return out


if __name__ == "__main__":
    main()
