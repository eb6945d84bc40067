"""Throughput of the FX-correlation hot path on one MI355X.

Default workload: 16 stations x 2 pols of 2-bit real samples at 256 MS/s
with three common tones, 8192-point FFTs -> 4096 channels, 31 250 FFTs (1 s
subintegration), FP64 delay model, visibilities of 120 baselines + 16 autos
x 4 products.  Synthetic input following the signal model is generated
directly in HBM (gcorr_amd.synth), so the timed region starts with the input
resident.

One step = one subintegration through FxKernel::process (gxc_process):
F-stage (unpack, fringe rotation, FFT, fractional-sample correction),
X-stage (cross-multiply-accumulate), finalise (normalise).

`value` is the FP32 path (the drop-in, fxkernel's precision).  The same
input then runs through the FP16-MFMA form (`--compare`), timed the same
way, and both forms are checked against the FP64 oracle on a sample of that
input (its first --accuracy-ffts FFTs, with the shapes the autotuner chose).

Prints one JSON line.  roofline: the dominant kernel's algorithmic bytes per
launch / its mean launch time (HIP events on the library's stream), against
the 8 TB/s HBM peak.  cpu_baseline: the oracle (CPU restatement of fxkernel)
timed on this host's cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Input Gsamples/s (all stations) + % HBM roofline; FP16-MFMA vs FP32 @1e-3 err"
HBM_PEAK_GBS = 8000.0
# dense MFMA peaks, TFLOP/s
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp16": 2500.0, "bf16": 2500.0, "fp8": 5000.0}
# worker threads the CPU baseline may use on a host shared with other GPUs'
# jobs (the share of one GPU)
CPU_THREADS_CAP = 16
ERR_BOUND = {"fp32": 1e-5, "fp16": 1e-3, "bf16": 8e-3, "fp8": 0.12}


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16", "fp8"],
                    help="precision of `value`")
    ap.add_argument("--compare", default="fp16",
                    help="comma-separated precisions timed beside `value` on the same input ('none': skip)")
    ap.add_argument("--accuracy-ffts", type=int, default=256,
                    help="FFTs of the input checked against the FP64 oracle (0: skip)")
    ap.add_argument("--numffts", type=int, default=0, help="override FFTs per subint (testing)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target wall time of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-h2d", action="store_true")
    ap.add_argument("--no-vdif", action="store_true", help="skip the VDIF ingest measurement")
    ap.add_argument("--overlap", action="store_true",
                    help="F of chunk k+1 on a second stream beside X of chunk k")
    ap.add_argument("--chunk-ffts", type=int, default=0, help="FFTs per F/X pass (0: automatic)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (the autotuner runs before warmup)")
    ap.add_argument("--shape", action="append", default=[], metavar="PREC:FBLOCK,XTILE,XSPLITS[,GRAPH]",
                    help="launch shapes for one precision instead of its autotuning (a profiled "
                         "run then traces only the kernels that are timed), e.g. fp32:256,8,8; "
                         "GRAPH 0 runs eagerly (default 1: graph replay)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (scales HBM counters)")
    return ap.parse_args(argv)


def given_shapes(specs):
    """--shape PREC:FBLOCK,XTILE,XSPLITS[,GRAPH] -> {prec: {"fblock", "xtile", "xsplits", "graph"}}."""
    out = {}
    for sp in specs:
        prec, _, vals = sp.partition(":")
        v = [int(x) for x in vals.split(",")]
        if prec not in ERR_BOUND or len(v) not in (3, 4):
            raise SystemExit("--shape wants PREC:FBLOCK,XTILE,XSPLITS[,GRAPH], got %r" % sp)
        out[prec] = {"fblock": v[0], "xtile": v[1], "xsplits": v[2], "graph": bool(v[3]) if len(v) == 4 else True}
    return out


def compare_list(spec, main):
    if not spec or spec == "none":
        return []
    out = []
    for p in spec.split(","):
        p = p.strip()
        if p not in ERR_BOUND:
            raise SystemExit("unknown precision in --compare: %r" % p)
        if p != main and p not in out:
            out.append(p)
    return out


def traffic_from_profiles(workload, precision, fblock, launches_per_step):
    """HBM bytes per launch of each stage measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE, scaled by a 1 GiB copy), committed as
    profiles/traffic_<workload>_<precision>.json for the F-stage workgroup
    size the autotuner chose there and the F-stage launches per step of that
    run; None when this run chose another shape or another chunking."""
    p = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (workload, precision))
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    if t.get("fblock") != fblock or t.get("fstage_launches_per_step") != launches_per_step:
        return None
    return t


def algorithmic_bytes(ws, precision, nrows):
    """(input bytes per sample, spectra bytes per sample, visibility bytes per step)."""
    w = ws[0]
    spec = {"fp32": 4.0, "fp16": 2.0, "bf16": 2.0, "fp8": 1.0}[precision]
    inb = 2.0 if w.is_complex else 0.25
    if w.is_complex:
        spec *= 2
    vis = sum(nrows * 4 * x.nchan * 8 for x in ws)
    return inb, spec, vis


def xstage_flops(ws):
    """Useful X-stage flops per step: 8 per complex MAC over M(M+1)/2 input pairs."""
    out = 0.0
    for x in ws:
        M = 2 * x.nant
        out += 8.0 * x.numffts * x.nchan * M * (M + 1) / 2
    return out


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": affinity}


def native_oracle():
    """The oracle rebuilt for this host's CPU (-march=native) in a temporary
    directory, or None (the committed portable build is then used)."""
    try:
        d = tempfile.mkdtemp(prefix="fxref_native_")
        r = subprocess.run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                            "NATIVE_DIR=" + d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        lib = os.path.join(d, "libfxref.so")
        if r.returncode == 0 and os.path.exists(lib):
            return lib, d
        shutil.rmtree(d, ignore_errors=True)
    except Exception:
        pass
    return None


def cpu_baseline(w, poly, offs, target_s):
    import numpy as np
    from oracle import fxref
    cpu = host_cpu()
    nthreads = max(1, min(cpu["affinity_cpus"], CPU_THREADS_CAP))
    built = native_oracle()
    flags = fxref.compile_flags(native=built is not None)
    if built:
        fxref.use_library(built[0])
    try:
        cfgf = lambda n: fxref.make_config(w.nant, w.nchan, n, w.lo, w.bw, nbit=w.nbit,
                                           is_complex=w.is_complex)
        # calibrate on one thread, then size each thread's subint to the target
        probe = max(8, min(w.numffts, 64))
        rng = np.random.default_rng(1)
        d = fxref.derive(cfgf(probe))
        data = [rng.integers(0, 256, d.bytes_per_station, dtype=np.uint8) for _ in range(w.nant)]
        t1 = fxref.bench(cfgf(probe), data, poly, offs, 1, 1)
        per_fft = t1 / probe
        n = int(max(probe, min(w.numffts, target_s / per_fft)))
        d = fxref.derive(cfgf(n))
        data = [rng.integers(0, 256, d.bytes_per_station, dtype=np.uint8) for _ in range(w.nant)]
        secs = fxref.bench(cfgf(n), data, poly, offs, nthreads, 1)
    finally:
        if built:
            fxref.use_library(None)
            shutil.rmtree(built[1], ignore_errors=True)
    samples = nthreads * w.nant * 2 * n * w.fftchannels
    single = w.nant * 2 * probe * w.fftchannels / t1
    return {
        "value": samples / secs / 1e9,
        "unit": "Gsamples/s",
        "cores": nthreads,
        "kind": "port",
        "sample": ("oracle FP32 fxkernel restatement (not IPP fxkernel; fxkernel is itself not "
                   "cache-blocked), %d threads x one subint of %d FFTs (%s shape), own kernel "
                   "instance and input copy per thread, as bench_fxkernel runs" % (nthreads, n, w.name)),
        "single_thread_value": single / 1e9,
        "seconds": secs,
        "cpu_model": cpu["model"],
        "host_logical_cpus": cpu["logical_cpus"],
        "host_affinity_cpus": cpu["affinity_cpus"],
        "threads_cap": CPU_THREADS_CAP,
        "compiler": "gcc " + flags,
    }


def make_kernel(gcorr_amd, ws, gens, precision, args, shape=None, numffts=None):
    """A handle for the workload (a Batch for the multi-subband one) with the
    generated device input and delays attached."""
    shape = shape or {}
    w = ws[0]
    if len(ws) > 1:
        k = gcorr_amd.Batch(w.nant, [(x.nchan, x.numffts, x.lo, x.bw) for x in ws],
                            precision=precision, autos=True, overlap=args.overlap)
        for sidx, (bufs, poly, offs) in enumerate(gens):
            k.set_input_device(sidx, [b.data_ptr() for b in bufs], keepalive=bufs)
            k.set_delays(sidx, poly, offs)
        return k
    bufs, poly, offs = gens[0]
    k = gcorr_amd.FxKernel(w.nant, w.nchan, numffts or w.numffts, w.nbit, w.lo, w.bw, w.is_complex,
                           precision=precision, autos=True, overlap=args.overlap,
                           chunk_ffts=args.chunk_ffts, **shape)
    k.setInputDevice([b.data_ptr() for b in bufs], keepalive=bufs)
    k.setDelays(poly, offs)
    return k


def time_precision(gcorr_amd, torch, ws, gens, precision, args):
    """Autotune (or take the --shape given), warm up and time `steps`
    subintegrations in `precision`; then the per-stage times from a profiled
    pass.  Returns (result dict, open handle)."""
    shape = given_shapes(args.shape).get(precision) if len(ws) == 1 else None
    k = make_kernel(gcorr_amd, ws, gens, precision, args, shape=shape)
    input_samples = sum(x.input_samples for x in ws)
    tuning = None
    if shape:
        tuning = dict(shape, given=True)
    elif not args.no_autotune:
        t = time.time()
        tuning = k.autotune()
        tuning["seconds"] = time.time() - t
        print("autotune %s: %s" % (precision, tuning), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        k.process()
    # Timed region: the steps as the library runs them by default (the tuned
    # launch mode, a captured graph unless eager measured faster).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k.process()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    print("timed %s: %.3f ms per step" % (precision, ms_per_step), file=sys.stderr, flush=True)
    # Stage times: the same steps again with HIP events around every launch
    # on the stream it runs on (this runs the launches eagerly).
    k.set_profiling(True)
    f_ms = x_ms = fin_ms = 0.0
    f_launch = x_launch = 0
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        k.process()
        tm = k.timings()
        f_ms += tm["fstage_ms"]
        x_ms += tm["xstage_ms"]
        fin_ms += tm["finalise_ms"]
        f_launch += tm["fstage_launches"]
        x_launch += tm["xstage_launches"]
    torch.cuda.synchronize()
    profiled_ms_per_step = (time.perf_counter() - t) * 1e3 / args.steps
    k.set_profiling(False)

    # algorithmic bytes per launch: input bits + nchan-bin spectra (F);
    # spectra read once + visibilities written (X)
    nrows = gcorr_amd.num_rows(ws[0].nant, autos=True)
    in_b, spec_b, vis_bytes = algorithmic_bytes(ws, precision, nrows)
    f_per_step = max(1, f_launch / args.steps)
    x_per_step = max(1, x_launch / args.steps)
    f_bytes = input_samples * (in_b + spec_b) / f_per_step
    x_bytes = (input_samples * spec_b + vis_bytes) / x_per_step
    f_avg = f_ms / max(1, f_launch)
    x_avg = x_ms / max(1, x_launch)
    xflops = xstage_flops(ws) / x_per_step
    x_tflops = xflops / (x_avg * 1e-3) / 1e12 if x_avg else None
    stages = {
        "fstage": {"ms_per_launch": f_avg, "launches_per_step": f_launch / args.steps,
                   "bytes_per_launch": f_bytes, "GBps": f_bytes / (f_avg * 1e-3) / 1e9 if f_avg else None},
        "xstage": {"ms_per_launch": x_avg, "launches_per_step": x_launch / args.steps,
                   "bytes_per_launch": x_bytes, "GBps": x_bytes / (x_avg * 1e-3) / 1e9 if x_avg else None,
                   "flops_per_launch": xflops, "TFLOPs": x_tflops,
                   "mfma_peak_TFLOPs": MFMA_PEAK_TFLOPS[precision],
                   "mfma_frac": x_tflops / MFMA_PEAK_TFLOPS[precision] if x_tflops else None},
        "finalise_ms": fin_ms / args.steps,
        "profiled_eager_ms_per_step": profiled_ms_per_step,
    }
    dom = "fstage" if f_ms >= x_ms else "xstage"
    achieved = stages[dom]["GBps"]
    traffic = traffic_from_profiles(args.workload, precision, (tuning or {}).get("fblock"),
                                    f_launch / args.steps)
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved else None,
        "traffic": (traffic or {}).get(dom),
        "algorithmic_bytes_per_launch": stages[dom]["bytes_per_launch"],
    }
    # what binds the profiled kernel, from its committed counters
    if traffic and traffic.get("limit", {}).get(dom):
        roofline["limit"] = traffic["limit"][dom]
    res = {"precision": precision, "ms_per_step": ms_per_step,
           "value": input_samples / (ms_per_step * 1e-3) / 1e9,
           "roofline": roofline, "stages": stages, "tuning": tuning}
    if precision == "fp32" and tuning and tuning.get("fblock") == -1:
        res["note"] = "FP32 F-stage on the matrix cores: operands as f16 hi + lo pairs (22 significant bits)"
    return res, k


def accuracy(gcorr_amd, ws, gens, results, args):
    """Each timed precision against the FP64 oracle on the first
    --accuracy-ffts FFTs of the same input: a handle over that many FFTs of
    the device buffers, with the launch shapes the autotuner chose."""
    import numpy as np
    from gcorr_amd import metrics
    from oracle import fxref
    w = ws[0]
    n = min(args.accuracy_ffts, w.numffts)
    bufs, poly, offs = gens[0]
    nbytes = gcorr_amd.station_bytes(w.nant, w.nchan, n, w.nbit, w.is_complex)
    host = [b[:nbytes].cpu().numpy() for b in bufs]
    cfg = fxref.make_config(w.nant, w.nchan, n, w.lo, w.bw, nbit=w.nbit, is_complex=w.is_complex,
                            autos=True)
    t = time.time()
    ref, rc = fxref.process(cfg, host, poly, offs, "f64")
    ref32, _ = fxref.process(cfg, host, poly, offs, "f32")
    out = {"sample": "FFTs 0..%d of the %d timed (the same device input, tuned launch shapes); "
                     "error = max over rows and products of ||V - V64|| / ||V64|| over channels"
                     % (n - 1, w.numffts),
           "ffts": n, "oracle_seconds": time.time() - t,
           "oracle_fp32_restatement": {"eps": metrics.eps(ref32, ref)}}
    for prec, r in results.items():
        shape = {}
        if r.get("tuning"):
            shape = {key: r["tuning"][key] for key in ("fblock", "xtile", "xsplits")}
        k = make_kernel(gcorr_amd, ws, gens, prec, args, shape=shape, numffts=n)
        k.process()
        v, c = k.visibilities(), k.row_counts()
        k.close()
        e = metrics.eps(v, ref)
        out[prec] = {"eps": e, "bound": ERR_BOUND[prec], "within_bound": bool(e <= ERR_BOUND[prec]),
                     "row_counts_equal": bool(np.array_equal(c, rc)),
                     "per_product": metrics.amp_phase_errors(v, ref)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build's path runs on a single MI355X "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import gcorr_amd
    from gcorr_amd import synth

    if args.workload == "c4":
        ws = synth.c4_subbands()
        if args.numffts:  # keep the subbands' 2:1 FFT-count pattern
            top = max(x.numffts for x in ws)
            ws = [x.scaled(max(8, x.numffts * args.numffts // top)) for x in ws]
    else:
        ws = [synth.WORKLOADS[args.workload]]
        if args.numffts:
            ws = [ws[0].scaled(args.numffts)]
    w = ws[0]
    dev = torch.device("cuda:0")
    t = time.time()
    gens = [synth.generate(x, device=dev) for x in ws]
    torch.cuda.synchronize()
    print("generated %s in %.1f s" % (args.workload, time.time() - t), file=sys.stderr, flush=True)

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    main_res, k = time_precision(gcorr_amd, torch, ws, gens, args.precision, args)
    tuning = main_res["tuning"]
    input_samples = sum(x.input_samples for x in ws)
    main_vis = k.visibilities() if len(ws) == 1 else None
    k.close()
    del k

    results = {args.precision: main_res}
    for prec in compare_list(args.compare, args.precision):
        r, kc = time_precision(gcorr_amd, torch, ws, gens, prec, args)
        kc.close()
        results[prec] = r
    comparison = {}
    for prec, r in results.items():
        if prec == args.precision:
            continue
        comparison[prec] = {
            "ms_per_step": r["ms_per_step"], "value": r["value"],
            "speedup_vs_%s" % args.precision: main_res["ms_per_step"] / r["ms_per_step"],
            "roofline": r["roofline"], "stages": r["stages"], "tuning": r["tuning"]}

    acc = None
    if args.accuracy_ffts > 0 and len(ws) == 1:
        acc = accuracy(gcorr_amd, ws, gens, results, args)
        print("accuracy: %s" % {p: acc[p]["eps"] for p in results}, file=sys.stderr, flush=True)

    shape = {} if tuning is None else {key: tuning[key] for key in ("fblock", "xtile", "xsplits")}
    # Input handed over in host memory (PCIe-inclusive; never `value`):
    # pageable buffers through gxc_process (copy, then compute), and pinned
    # buffers streamed with gxc_submit (copy of subint k+1 under compute of k).
    pcie = None
    if not args.no_h2d and len(ws) == 1:
        bufs, poly, offs = gens[0]
        host = [b.cpu().numpy() for b in bufs]
        kh = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, w.is_complex,
                                precision=args.precision, autos=True, **shape)
        kh.setInputData(host)
        kh.setDelays(poly, offs)
        kh.process()
        n = 3
        t = time.perf_counter()
        for _ in range(n):
            kh.process()
        pcie = {"pageable_sync_gsamples_per_s": w.input_samples * n / (time.perf_counter() - t) / 1e9}
        if 2 * w.nant * w.station_bytes <= (8 << 30):
            pins = [[gcorr_amd.PinnedBuffer(w.station_bytes) for _ in range(w.nant)] for _ in range(2)]
            for row in pins:
                for pb, hb in zip(row, host):
                    pb.array[:] = hb
            kh.integrate_begin()
            for i in range(2):
                kh.submit([pb.array for pb in pins[i & 1]], poly, offs)
            kh.integrate_end()
            n = max(4, min(args.steps, 10))
            kh.integrate_begin()
            t = time.perf_counter()
            for i in range(n):
                kh.submit([pb.array for pb in pins[i & 1]], poly, offs)
            kh.integrate_end()
            dt = time.perf_counter() - t
            pcie["streamed_pinned_gsamples_per_s"] = w.input_samples * n / dt / 1e9
            pcie["streamed_h2d_GBps"] = w.nant * w.station_bytes * n / dt / 1e9
            pcie["streamed_subints"] = n
            for row in pins:
                for pb in row:
                    pb.free()
        kh.close()
        del host
        print("pcie: %s" % pcie, file=sys.stderr, flush=True)

    # VDIF ingest (gxc_set_input_vdif): the same input as 8000-byte-payload
    # frames in HBM, placed by the GPU, then correlated; wall-clock per
    # subintegration, every station ingested once.
    vd = None
    if not args.no_vdif and len(ws) == 1 and w.nant * w.station_bytes <= (4 << 30):
        from gcorr_amd import vdif
        bufs, poly, offs = gens[0]
        payload = 8000
        spf = payload * 8 // (w.nbit * 2 * (2 if w.is_complex else 1))
        fps = max(1, int(round(w.fftchannels / spf * w.numffts)))
        frames = [torch.from_numpy(vdif.frames_from_station(b.cpu().numpy(), payload, w.nbit, w.is_complex,
                                                            frames_per_second=fps).reshape(-1)).to(dev)
                  for b in bufs]
        kv = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, w.is_complex,
                                precision=args.precision, autos=True, **shape)
        kv.setDelays(poly, offs)

        def ingest():
            for a, f in enumerate(frames):
                kv.setInputVDIF(a, f.data_ptr(), f.numel(), frames_per_second=fps,
                                frame_bytes=payload + 32, header_bytes=32, stats=False)

        ingest()
        kv.process()
        same = bool(np.array_equal(kv.visibilities(), main_vis))
        n = 5
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            ingest()
        kv.synchronize()
        ing = (time.perf_counter() - t) / n
        t = time.perf_counter()
        for _ in range(n):
            ingest()
            kv.process()
        both = (time.perf_counter() - t) / n
        fbytes = sum(f.numel() for f in frames)
        vd = {"payload_bytes": payload, "frame_bytes": payload + 32, "ingest_ms": ing * 1e3,
              "ingest_GBps": (fbytes + w.nant * w.station_bytes) / ing / 1e9,
              "with_process_gsamples_per_s": input_samples / both / 1e9,
              "visibilities_equal_device_input": same}
        kv.close()
        del frames
        print("vdif: %s" % vd, file=sys.stderr, flush=True)

    # Requantiser (gxc_requantise_2bit): one station's float voltages of this
    # workload (both pols) -> its packed 2-bit buffer, HIP events on the
    # stream it runs on (torch's current stream).
    rq = None
    if not args.no_vdif and not w.is_complex and len(ws) == 1:
        n = 2 * w.station_bytes  # time samples per pol
        x = torch.randn(2, n, device=dev)
        out = torch.empty(w.station_bytes, dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream()
        for _ in range(2):
            gcorr_amd.requantise_2bit(x[0].data_ptr(), x[1].data_ptr(), n, 0.98, out.data_ptr(),
                                      stream.cuda_stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record(stream)
        for _ in range(reps):
            gcorr_amd.requantise_2bit(x[0].data_ptr(), x[1].data_ptr(), n, 0.98, out.data_ptr(),
                                      stream.cuda_stream)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 2 * n * 4 + w.station_bytes
        rq = {"samples_per_pol": n, "ms": ms, "GBps": nbytes / (ms * 1e-3) / 1e9,
              "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        del x, out

    cpu = None
    if not args.no_cpu_baseline and len(ws) == 1:
        bufs, poly, offs = gens[0]
        cpu = cpu_baseline(w, poly, offs, args.cpu_seconds)
        print("cpu baseline: %.3f Gsamples/s on %d threads" % (cpu["value"], cpu["cores"]),
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": main_res["value"],
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic (seeded signal model generated in HBM: common signal, quadratic delays, fringes, 2-bit quantised)",
        "config": {
            "workload": w.name if len(ws) == 1 else "8st_2pol_8subbands_16-128MHz_256-4096ch_0.5s",
            "stations": w.nant, "pols": 2,
            "nchan": [x.nchan for x in ws] if len(ws) > 1 else w.nchan,
            "fft": [x.fftchannels for x in ws] if len(ws) > 1 else w.fftchannels,
            "numffts": [x.numffts for x in ws] if len(ws) > 1 else w.numffts,
            "bandwidth_hz": [x.bw for x in ws] if len(ws) > 1 else w.bw, "lo_hz": w.lo,
            "input": "complex int8" if w.is_complex else "real 2-bit",
            "precision": args.precision, "autos": True,
            "overlap": args.overlap, "chunk_ffts": args.chunk_ffts or "auto",
            "input_samples_per_step": input_samples,
        },
        "roofline": main_res["roofline"],
        "stages": main_res["stages"],
        "tuning": tuning,
        "compare": comparison,
        "accuracy_vs_fp64_oracle": acc,
        "gbps_input": main_res["value"] * (16 if w.is_complex else 2),
        "pcie_inclusive": pcie,
        "vdif": vd,
        "requantise_2bit": rq,
        "cpu_baseline": cpu,
    }
    if "note" in main_res:
        out["note"] = main_res["note"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
