"""Benchmark: HiCOO MTTKRP on one MI355X.

Workload: order-3 tensor 2^20 x 2^20 x 2^16, 100M unique nonzeros, each
mode's index Zipf(1.1) over its dimension and randomly relabelled, values
U(0,1); HiCOO with sb=7, sk=14, sc=14; rank 32 factors U(0,1).  A step is
one MTTKRP per mode (modes 0, 1, 2) on the device-resident tensor, i.e. one
pass of the hot path over the whole tensor.  Inputs are in HBM before the
timed region starts.

Printed JSON (one line): value = milliseconds per step (all three modes,
FP32 factors, lower is better); the roofline of the HiCOO kernel with
algorithmic bytes per launch
    B = nnz*(N*1 + 4) + nb*(N*4 + 8) + sum_m ndims[m]*R*4
(the reference's convention, benchmark/tensor/mttkrp.c:304-307, applied to
the HiCOO arrays) divided by the measured average duration of one mode's
MTTKRP (HIP events on the launch stream; this includes the output zero
fill), against 8 TB/s, and beside it the attainable bandwidth of a 1 GiB
device copy measured in the same run; the BF16-factor form's time; and the
CPU baseline on a bounded sample of the same tensor.  roofline.traffic is
the HBM bytes per launch from the committed rocprofv3 counter summary of
this kernel, used only when that summary names the current kernel source
(its SHA-256), else null.

Every leg states its accuracy: "err_vs_fp64" holds, per mode, the relative
Frobenius error of the GPU output on the 256 heaviest output rows plus 256
random non-empty rows against the oracle's FP64 MTTKRP of those rows (BF16:
against FP64 on the same bf16-rounded factors), computed after the timed
regions.

Other legs (each skippable):
  "coo"        ptiMTTKRP's device path on the same tensor and factors;
  "precision"  the precision form: order-4 10M x 50k x 2k x 100, 200M
               nonzeros (Zipf(1.2) on modes 0-1, uniform on 2-3, lognormal
               values), R = 64, FP32 versus BF16 factors per mode;
  "cpd"        CP-ALS on an order-5 1M x 500k x 100k x 1k x 24 tensor with
               150M nonzeros, R = 16, 20 iterations: all-FP32 versus the
               BF16-then-FP32 schedule (s/iter, final fits), the MTTKRP /
               dense-step split, and the host CP-ALS restatement of
               ptiOmpCpdAls on a sample;
  "completion" the CP-completion loss (autograd forward + backward) on the
               50M-entry order-3 tensor, with its FP64 loss and gradient
               rows checked.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-cpd] [--no-coo]
                    [--no-precision] [--no-completion]
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--nnz", type=float, default=100e6)
    p.add_argument("--rank", type=int, default=32)
    p.add_argument("--seed", type=int, default=42)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-bf16", action="store_true")
    p.add_argument("--cpu-sample-nnz", type=float, default=8e6)
    p.add_argument("--no-cpd", action="store_true", help="skip the CPD-ALS leg")
    p.add_argument("--no-coo", action="store_true", help="skip the COO (ptiMTTKRP) leg")
    p.add_argument("--no-completion", action="store_true", help="skip the CP-completion (autograd) leg")
    p.add_argument("--no-precision", action="store_true", help="skip the precision-form (order-4, R=64) leg")
    p.add_argument("--no-errors", action="store_true", help="skip the FP64 accuracy checks")
    p.add_argument("--prec-nnz", type=float, default=200e6)
    p.add_argument("--prec-cpu-sample-nnz", type=float, default=1e6,
                   help="nonzeros of the precision tensor the CPU baseline runs on (mostly one thread)")
    p.add_argument("--cpd-cpu-stride", type=int, default=15,
                   help="host CP-ALS baseline runs on every k-th nonzero of the CPD tensor")
    p.add_argument("--cmp-nnz", type=float, default=50e6)
    p.add_argument("--cpd-nnz", type=float, default=150e6)
    p.add_argument("--cpd-rank", type=int, default=16)
    p.add_argument("--cpd-iters", type=int, default=20)
    p.add_argument("--cpd-bf16-iters", type=int, default=10)
    p.add_argument("--copy-calibration", type=float, default=0.0,
                   help="only copy a GiB-sized float buffer 3 times (HBM counter calibration) and exit")
    return p.parse_args()


def copy_calibration(gib):
    """Plain device-to-device copies of a known size, so that rocprofv3's
    FETCH_SIZE / WRITE_SIZE readings can be scaled to bytes."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_copy": n * 4, "copies": 3}), flush=True)


KERNEL = "hicoo_stream_kernel<3, 4, 8, float, 8, 0, false>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hicoo_stream_kernel_zipf100M_R32_counters.md")


# The kernel's source and the code that lays out the stream forms it reads:
# a change to either can move the traffic.
KERNEL_SOURCES = [os.path.join(ROOT, "hiparti_amd", "csrc", "device", f)
                  for f in ("mttkrp_kernels.hpp", "launch.hpp", "hicoo_dev.hip", "coo_dev.hip")]


def kernel_source_sha():
    """SHA-256 over the concatenated KERNEL_SOURCES."""
    import hashlib
    h = hashlib.sha256()
    for path in KERNEL_SOURCES:
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(which="traffic"):
    """HBM bytes per launch of the stream kernel on this workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE,
    calibrated on a plain copy), when it was taken on the current kernel:
    same kernel name and the same SHA-256 of the kernel and stream-form
    sources (KERNEL_SOURCES).  which:
    "traffic" (FP32 HiCOO), "bf16_traffic" or "coo_traffic"."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    kern = re.search(r"^kernel: (.*)$", text, re.M)
    sha = re.search(r"^kernel_source_sha256: ([0-9a-f]+)$", text, re.M)
    val = re.search(r"^%s_bytes_per_launch: (\d+)$" % which, text, re.M)
    if not kern or not val or not sha or kern.group(1).strip() != KERNEL or sha.group(1) != kernel_source_sha():
        return None
    return int(val.group(1))


def attainable_bandwidth(gib=1.0, reps=5):
    """GB/s of a device-to-device copy of `gib` GiB (read + write bytes over
    the hipEvent-timed duration, best of `reps`): the bandwidth HBM gives a
    plain streaming kernel on this box, reported beside the 8 TB/s spec."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = None
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b)
        best = ms if best is None else min(best, ms)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * n * 4 / (best * 1e-3) / 1e9


def host_facts(nthreads):
    """The CPU the baseline ran on: model name, logical CPUs, threads used,
    OMP_PROC_BIND."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "threads": nthreads,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "unset (no binding)")}


# ---------------------------------------------------------------------------
# Accuracy against the FP64 oracle (run after the timed regions).
# ---------------------------------------------------------------------------
def sample_rows(idx, nrows, seed, n_heavy=256, n_random=256):
    """The n_heavy output rows with the most nonzeros and n_random seeded
    random non-empty rows among the rest."""
    counts = np.bincount(idx, minlength=nrows)
    heavy = np.argsort(counts, kind="stable")[::-1][:n_heavy]
    rest = np.setdiff1d(np.flatnonzero(counts), heavy)
    rng = np.random.default_rng(seed)
    rand = rng.choice(rest, size=min(n_random, len(rest)), replace=False)
    return np.concatenate([heavy, rand]).astype(np.uint32)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a))


def fp64_rows(inds, vals, facs, mode, rows, nthreads):
    """The oracle's FP64 MTTKRP of the given output rows (the checker)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O.mttkrp_rows(inds, vals, facs, mode, rows, nthreads=nthreads)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# The oracle's OpenMP restatements against the reference's own ptiOmpMTTKRP /
# ptiOmpMTTKRPHiCOO, both built from source and timed on the same 2M-nonzero
# sample of the 100M-nonzero tensor on 8 host cores (ms per 3-mode step,
# extrapolated): restatement / reference.
PORT_VS_REFERENCE = {"hicoo": 1.03, "coo": 1.02}


def _sample_prefix(h, sample_nnz):
    """The leading blocks holding about sample_nnz nonzeros, as superblocks:
    whole superblocks while they fit, then the first blocks of the next one
    (a superblock can hold most of a tensor whose first mode is Zipf and not
    relabelled).  Returns (kptr of the sample, b_end blocks, z_end nonzeros,
    whether the last superblock was cut)."""
    a = h.address
    nk_total = int(a.kptr.len) - 1
    kptr = np.ctypeslib.as_array(a.kptr.data, shape=(nk_total + 1,))
    bptr_all = np.ctypeslib.as_array(a.bptr.data, shape=(int(a.bptr.len),))
    k_full = int(np.searchsorted(bptr_all[kptr], sample_nnz, side="right")) - 1   # superblocks that fit whole
    k_full = max(0, min(k_full, nk_total))
    sk = list(kptr[:k_full + 1])
    cut = False
    if k_full < nk_total and bptr_all[kptr[k_full]] < sample_nnz:
        nb_total = int(kptr[nk_total])
        b_end = int(np.searchsorted(bptr_all, sample_nnz))
        b_end = max(int(kptr[k_full]) + 1, min(b_end, int(kptr[k_full + 1]), nb_total))
        cut = b_end < int(kptr[k_full + 1])
        sk.append(b_end)
    b_end = int(sk[-1])
    return np.array(sk, np.uint64), b_end, int(bptr_all[b_end]), cut


def _time_modes(run, N, what="cpu baseline"):
    """One warm-up and two timed calls per mode; seconds per N-mode pass."""
    total = 0.0
    for mode in range(N):
        run(mode)
        t0 = time.perf_counter()
        for _ in range(2):
            run(mode)
        dt = (time.perf_counter() - t0) / 2
        log("%s: mode %d %.3f s per call" % (what, mode, dt))
        total += dt
    return total


def cpu_baseline(h, ndims, facs_np, R, sample_nnz, nthreads):
    """The CPU baseline on the leading superblocks holding about sample_nnz
    nonzeros of the same HiCOO tensor, one warm-up and two timed calls per
    mode, extrapolated per nonzero to the whole tensor (ms per N-mode step).
    It is the oracle's restatement of the reference's ptiOmpMTTKRPHiCOO with
    tk = host threads and tb = 1 (superblock-parallel with per-element
    atomics, src/sptensor/hicoo/mttkrp_omp.c:184-204,400-467); the reference
    library itself is built only where its sources are, to pin that
    restatement (PORT_VS_REFERENCE).  Returns (ms, kind, z_end, a text naming
    the sample)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    a = h.address
    N = len(ndims)
    skptr, b_end, z_end, cut = _sample_prefix(h, sample_nnz)
    sub = {
        "kptr": skptr,
        "bptr": np.ctypeslib.as_array(a.bptr.data, shape=(b_end + 1,)).copy(),
        "binds": np.stack([np.ctypeslib.as_array(a.binds[m].data, shape=(b_end,)) for m in range(N)]),
        "einds": np.stack([np.ctypeslib.as_array(a.einds[m].data, shape=(z_end,)) for m in range(N)]),
        "values": np.ctypeslib.as_array(a.values.data, shape=(z_end,)).copy(),
    }

    def run(mode):
        O.mttkrp_hicoo(sub, int(a.sb_bits), facs_np, mode, kind="omp", nthreads=nthreads)
    total = _time_modes(run, N, "HiCOO cpu baseline")
    nk = len(skptr) - 1
    if cut:
        text = ("%d whole superblocks and the first %d blocks of the next (%d of its %d blocks)"
                % (nk - 1, b_end - int(skptr[-2]), b_end - int(skptr[-2]),
                   int(np.ctypeslib.as_array(a.kptr.data, shape=(int(a.kptr.len),))[nk] - skptr[-2])))
    else:
        text = "the first %d superblocks" % nk
    return total / z_end * int(a.nnz) * 1e3, "port", z_end, text


def coo_leg(X, ndims, dfacs, facs_np, outs, R, args, nthreads, rows=None):
    """ptiMTTKRP on the device (COO, per-mode copies sorted by the output
    index, cut into 128-row blocks and run by the stream kernel) on the same
    tensor and factors: ms per three-mode step with HIP events,
    its roofline with the reference's COO convention B = nnz*(N*4 + 4) +
    sum_m ndims[m]*R*4 (benchmark/tensor/mttkrp.c:304-307), and the CPU
    baseline: the oracle's restatement of ptiOmpMTTKRP (src/sptensor/
    mttkrp_omp.c:125-252, nonzero-parallel with atomics) on the first
    cpu_sample_nnz nonzeros, extrapolated per nonzero."""
    import torch
    from hiparti_amd.device import DevCOO
    N = len(ndims)
    nnz = X.nnz()
    t0 = time.time()
    Dc = DevCOO(X)
    for m in range(N):
        Dc.prepare(m)
    torch.cuda.synchronize()
    log("COO upload and per-mode sorts in %.1f s" % (time.time() - t0))

    def step(events=None):
        for mode in range(N):
            if events is not None:
                events[mode][0].record()
            Dc.mttkrp(dfacs, mode, R=R, out=outs[mode])
            if events is not None:
                events[mode][1].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(N)]
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events=ev[k])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    per_mode = [float(np.mean([ev[k][m][0].elapsed_time(ev[k][m][1]) for k in range(args.steps)])) for m in range(N)]
    # The outputs' sampled rows, for the accuracy check after the timed legs.
    got_rows = None
    if rows is not None:
        got_rows = []
        for mode in range(N):
            Dc.mttkrp(dfacs, mode, R=R, out=outs[mode])
            got_rows.append(outs[mode][torch.from_numpy(rows[mode].astype(np.int64)).cuda()].cpu().numpy())
    del Dc
    torch.cuda.empty_cache()
    alg = nnz * (N * 4 + 4) + sum(d * R * 4 for d in ndims)
    achieved = alg / (np.mean(per_mode) * 1e-3) / 1e9
    res = {"ms_per_step": round(ms, 4), "per_mode_ms": [round(x, 4) for x in per_mode],
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("coo_traffic"),
                        "kernel": "hicoo_stream_kernel<3, 4, 8, float, 8, 0, false> on the COO copy's 128-row blocks "
                                  "(+ output zero fill)",
                        "alg_bytes_per_launch": int(alg)}}
    if not args.no_cpu:
        t0 = time.time()
        z = int(min(args.cpu_sample_nnz, nnz))
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        inds = X.inds()[:, :z].copy()
        vals = X.values()[:z].copy()

        def run(mode):
            O.mttkrp(inds, vals, facs_np, mode, kind="omp", nthreads=nthreads)
        kind, who = "port", "oracle restatement of ptiOmpMTTKRP"
        total = _time_modes(run, N)
        res["cpu_baseline"] = {"value": round(total / z * nnz * 1e3, 1),
                               "unit": "ms per step (3-mode COO MTTKRP, extrapolated per nonzero)",
                               "cores": nthreads, "kind": kind,
                               "sample": "%s (tk=%d) on the first %d nonzeros of the same tensor, all 3 modes, "
                                         "1 warm-up + 2 timed calls per mode" % (who, nthreads, z)}
        res["cpu_baseline"].update(host_facts(nthreads))
        res["cpu_baseline"]["port_vs_reference"] = PORT_VS_REFERENCE["coo"]
        log("coo cpu baseline %.1f s" % (time.time() - t0))
    return {"coo": res}, got_rows


def time_per_mode(D, facs, outs, R, N, steps, warmup, bf16=False):
    """Per-mode MTTKRP milliseconds (HIP events on the launch stream, mean
    over `steps` steps after `warmup`) and wall milliseconds per N-mode step."""
    import torch

    def step(events=None):
        for mode in range(N):
            if events is not None:
                events[mode][0].record()
            if bf16:
                D.mttkrp_bf16(facs, mode, R=R, out=outs[mode])
            else:
                D.mttkrp(facs, mode, R=R, out=outs[mode])
            if events is not None:
                events[mode][1].record()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(N)]
          for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        step(events=ev[k])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    per_mode = [float(np.mean([ev[k][m][0].elapsed_time(ev[k][m][1]) for k in range(steps)])) for m in range(N)]
    return per_mode, wall


def hicoo_bytes(nnz, nb, ndims, R, mode=None, bf16=False):
    """Algorithmic bytes of one HiCOO MTTKRP (the reference's convention on
    the HiCOO arrays): nnz*(N + 4) + nb*(4N + 8) + sum_m ndims[m]*R*s_m, with
    s_m = 4 (FP32), or 2 for the gathered BF16 factors and 4 for the FP32
    output (mode `mode`)."""
    N = len(ndims)
    fac = sum(d * R * (2 if (bf16 and m != mode) else 4) for m, d in enumerate(ndims))
    return nnz * (N + 4) + nb * (4 * N + 8) + fac


def precision_leg(args, nthreads):
    """The precision form: order-4 HiCOO tensor 10M x 50k x 2k x 100 with
    200M nonzeros (Zipf(1.2) on modes 0-1, unpermuted; uniform on modes 2-3;
    lognormal(0,1) values), HiCOO sb=7 sk=14 sc=14, R = 64 factors U(0,1).
    Per mode: FP32 versus BF16-factor MTTKRP time (HIP events), the HiCOO
    roofline of each (s_m = 2 for the gathered BF16 factors), the error of
    each against the oracle's FP64 rows (BF16: on the rounded factors), and
    the oracle's restatement of ptiOmpMTTKRPHiCOO on a sample."""
    import torch
    from hiparti_amd import COOTensor, DenseMatrix
    from hiparti_amd.device import DevHiCOO
    ndims = [10_000_000, 50_000, 2_000, 100]
    N, R = 4, 64
    nnz = int(args.prec_nnz)
    t0 = time.time()
    X = COOTensor.synthetic(ndims, nnz, dist=[1, 1, 0, 0], zipf_s=[1.2, 1.2, 1, 1], permute=[0] * 4,
                            lognormal=True, seed=args.seed + 2, nthreads=nthreads)
    inds = vals = None
    if not args.no_errors:
        inds, vals = X.inds(), X.values()
    H = X.toHiCOO(7, 14, 14, nthreads=nthreads, device=True)
    nb = H.nblocks()
    X.free()
    D = DevHiCOO(H)
    torch.cuda.synchronize()
    log("precision tensor: %d nnz, nb=%d, ready in %.1f s (%.1f GB on the device)"
        % (nnz, nb, time.time() - t0, D.bytes / 1e9))
    facs_np = []
    for m, d in enumerate(ndims):
        dm = DenseMatrix(d, R)
        dm.uniform(args.seed * 100 + 10 + m)
        facs_np.append(dm.numpy())
    dfacs = [torch.from_numpy(f).cuda() for f in facs_np]
    bfacs = [f.to(torch.bfloat16) for f in dfacs]
    outs = [torch.empty((d, R), dtype=torch.float32, device="cuda") for d in ndims]
    t0 = time.perf_counter()
    for mode in range(N):   # first calls build the stream forms (setup, not timed)
        D.mttkrp(dfacs, mode, R=R, out=outs[mode])
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    p32, w32 = time_per_mode(D, dfacs, outs, R, N, args.steps, max(args.warmup, 1))
    rows = got32 = got16 = None
    if not args.no_errors:
        rows = [sample_rows(inds[m], ndims[m], seed=30 + m) for m in range(N)]
        got32 = [outs[m][torch.from_numpy(rows[m].astype(np.int64)).cuda()].cpu().numpy() for m in range(N)]
    p16, w16 = time_per_mode(D, bfacs, outs, R, N, args.steps, max(args.warmup, 1), bf16=True)
    if not args.no_errors:
        got16 = [outs[m][torch.from_numpy(rows[m].astype(np.int64)).cuda()].cpu().numpy() for m in range(N)]
    del D, dfacs, bfacs, outs
    torch.cuda.empty_cache()
    b32 = [hicoo_bytes(nnz, nb, ndims, R) for m in range(N)]
    b16 = [hicoo_bytes(nnz, nb, ndims, R, mode=m, bf16=True) for m in range(N)]
    res = {"workload": "HiCOO MTTKRP all 4 modes, 10M x 50k x 2k x 100, %d nnz, Zipf(1.2) modes 0-1, uniform modes "
                       "2-3, lognormal values, sb=7 sk=14 sc=14, R=%d" % (nnz, R),
           "nblocks": nb, "stream_build_s": round(setup_s, 3),
           "fp32_ms_per_step": round(w32, 4), "fp32_per_mode_ms": [round(x, 4) for x in p32],
           "bf16_ms_per_step": round(w16, 4), "bf16_per_mode_ms": [round(x, 4) for x in p16],
           "fp32_hbm_frac_per_mode": [round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for b, t in zip(b32, p32)],
           "bf16_hbm_frac_per_mode": [round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for b, t in zip(b16, p16)],
           "alg_bytes_per_launch_fp32": int(b32[0]), "alg_bytes_per_launch_bf16": [int(b) for b in b16]}
    log("precision: FP32 %s ms, BF16 %s ms per mode" % (res["fp32_per_mode_ms"], res["bf16_per_mode_ms"]))
    if not args.no_errors:
        t0 = time.time()
        rounded = [torch.from_numpy(f).to(torch.bfloat16).float().numpy() for f in facs_np]
        e32, e16 = [], []
        for m in range(N):
            e32.append(rel_err(got32[m][:, :R], fp64_rows(inds, vals, facs_np, m, rows[m], nthreads)))
            e16.append(rel_err(got16[m][:, :R], fp64_rows(inds, vals, rounded, m, rows[m], nthreads)))
        res["err_vs_fp64"] = {"fp32": [float("%.3e" % e) for e in e32], "bf16": [float("%.3e" % e) for e in e16],
                              "rows": "per mode: the 256 heaviest output rows + 256 random non-empty rows; BF16 "
                                      "against FP64 on the bf16-rounded factors"}
        log("precision errors in %.1f s: fp32 %s bf16 %s" % (time.time() - t0, e32, e16))
    if not args.no_cpu:
        t0 = time.time()
        cpu_ms, kind, z_s, where = cpu_baseline(H, ndims, facs_np, R, args.prec_cpu_sample_nnz, nthreads)
        who = "oracle restatement of ptiOmpMTTKRPHiCOO"
        res["cpu_baseline"] = {"value": round(cpu_ms, 1),
                               "unit": "ms per step (4-mode FP32 HiCOO MTTKRP, extrapolated per nonzero)",
                               "cores": nthreads, "kind": kind,
                               "sample": "%s (tk=%d, tb=1) on %s = %d nonzeros of the same tensor, all 4 modes, "
                                         "1 warm-up + 2 timed calls per mode; the first superblock holds most of "
                                         "the nonzeros (mode 0 is Zipf(1.2), not relabelled), and the reference's "
                                         "superblock-parallel schedule runs it on one thread"
                                         % (who, nthreads, where, z_s)}
        res["cpu_baseline"].update(host_facts(nthreads))
        res["cpu_baseline"]["port_vs_reference"] = PORT_VS_REFERENCE["hicoo"]
        log("precision cpu baseline %.1f s" % (time.time() - t0))
    H.free()
    return {"precision": res}


def cpd_leg(args, nthreads):
    """CP-ALS on the device: order-5 tensor 1M x 500k x 100k x 1k x 24,
    Zipf(1.1) indices on modes 0-2 and uniform on modes 3-4, values U(0,1),
    HiCOO sb=7 sk=14 sc=14, rank 16, a fixed number of iterations from the
    same seeded U(0,1) factors: all-FP32 MTTKRPs, then the schedule whose
    first iterations gather BF16 factor copies.  Returns s/iter and the final
    fits of both runs, the MTTKRP share of an iteration (the five MTTKRPs
    timed alone with HIP events) and its HiCOO roofline, and the host CP-ALS
    restatement of ptiOmpCpdAls timed on every k-th nonzero."""
    import torch
    from hiparti_amd import COOTensor
    from hiparti_amd import cpd as CP
    from hiparti_amd.device import DevHiCOO
    ndims = [1000000, 500000, 100000, 1000, 24]
    N = 5
    nnz = int(args.cpd_nnz)
    R = args.cpd_rank
    t0 = time.time()
    X = COOTensor.synthetic(ndims, nnz, dist=[1, 1, 1, 0, 0], zipf_s=[1.1, 1.1, 1.1, 0, 0],
                            permute=[0, 0, 0, 0, 0], seed=args.seed + 4, nthreads=nthreads)
    sample = None
    if not args.no_cpu:
        k = max(args.cpd_cpu_stride, 1)
        sample = (np.ascontiguousarray(X.inds()[:, ::k]), np.ascontiguousarray(X.values()[::k]), k)
    H = X.toHiCOO(7, 14, 14, nthreads=nthreads, device=True)
    X.free()
    D = DevHiCOO(H)
    nb = H.nblocks()
    H.free()
    torch.cuda.synchronize()
    log("cpd tensor: %d nnz, nb=%d, ready in %.1f s" % (nnz, nb, time.time() - t0))
    init = CP.init_factors(ndims, R, "uniform", seed=args.seed)
    # One warm-up iteration, then the timed runs from the same factors.
    CP.cp_als(D, R, niters=1, tol=0.0, factors=[f.clone() for f in init])
    runs = {}
    last = None
    for name, bf in (("fp32", 0), ("bf16_fp32", args.cpd_bf16_iters)):
        facs = [f.clone() for f in init]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = CP.cp_als(D, R, niters=args.cpd_iters, tol=0.0, factors=facs, bf16_iters=bf)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        runs[name] = (sec / max(res.iters, 1), res.fit, res.iters)
        last = facs
        log("cpd %s: %.4f s/iter, fit %.6f after %d iterations" % (name, runs[name][0], res.fit, res.iters))
    # The five MTTKRPs of an iteration on their own (FP32 and BF16 factors).
    outs = [torch.empty((d, R), dtype=torch.float32, device="cuda") for d in ndims]
    m32, _ = time_per_mode(D, last, outs, R, N, 5, 1)
    bl = [f.to(torch.bfloat16) for f in last]
    m16, _ = time_per_mode(D, bl, outs, R, N, 5, 1, bf16=True)
    del D, outs, bl
    torch.cuda.empty_cache()
    it_bytes = sum(hicoo_bytes(nnz, nb, ndims, R) for _ in range(N))
    mt_ms = sum(m32)
    out = {"workload": "CP-ALS order-5 1M x 500k x 100k x 1k x 24, %d nnz, Zipf(1.1) modes 0-2, uniform modes 3-4, "
                       "R=%d, %d iterations, HiCOO sb=7 sk=14" % (nnz, R, args.cpd_iters),
           "nblocks": nb,
           "fp32_s_per_iter": round(runs["fp32"][0], 5), "fp32_fit": runs["fp32"][1],
           "bf16_iters": args.cpd_bf16_iters,
           "bf16_fp32_s_per_iter": round(runs["bf16_fp32"][0], 5), "bf16_fp32_fit": runs["bf16_fp32"][1],
           "fit_delta": abs(runs["fp32"][1] - runs["bf16_fp32"][1]),
           "mttkrp_ms_per_iter_fp32": round(mt_ms, 3), "mttkrp_per_mode_ms_fp32": [round(x, 3) for x in m32],
           "mttkrp_ms_per_iter_bf16": round(sum(m16), 3),
           "dense_and_other_ms_per_iter_fp32": round(runs["fp32"][0] * 1e3 - mt_ms, 3),
           "mttkrp_roofline": {"alg_bytes_per_iter": int(it_bytes),
                               "achieved_GBps": round(it_bytes / (mt_ms * 1e-3) / 1e9, 1),
                               "frac": round(it_bytes / (mt_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    if sample is not None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        si, sv, k = sample
        init_np = [f.cpu().numpy() for f in init]
        t0 = time.time()
        iters = 2
        _, _, fits, tm, to = O.omp_cpd_als(si, sv, ndims, init_np, iters, 0.0, nthreads=nthreads)
        per_iter = (tm * k + to) / iters
        out["cpu_baseline"] = {"value": round(per_iter, 3), "unit": "s per iteration (extrapolated: MTTKRP time x %d, "
                                                                     "dense step as measured)" % k,
                               "cores": nthreads, "kind": "port",
                               "sample": "oracle restatement of ptiOmpCpdAls (src/sptensor/cpd_omp.c:32-244, "
                                         "use_reduce=0; FP64 Cholesky-Banachiewicz solve) on every %d-th nonzero "
                                         "(%d nonzeros) of the same tensor with the full factor sizes, %d iterations "
                                         "from the same initial factors: %.3f s in MTTKRPs, %.3f s elsewhere"
                                         % (k, sv.size, iters, tm, to),
                               "sample_s_per_iter": round((tm + to) / iters, 3)}
        out["cpu_baseline"].update(host_facts(nthreads))
        log("cpd cpu baseline (port) %.1f s" % (time.time() - t0))
    return {"cpd": out}


def completion_leg(args):
    """CP-completion loss on the device: order-3 tensor 500k x 500k x 10k with
    uniform indices, observed values from a planted rank-8 model plus N(0,
    0.01) noise, rank 32, factors 0.5*U(0,1).  A step is the autograd forward
    plus backward of the loss with every factor requiring a gradient: the
    loss and the three gradient MTTKRPs of the residuals (residuals formed
    inside each MTTKRP), then their scaling.  Also times the forward-only
    kernel (sampled reconstruction, residuals and FP64 loss), with its
    bandwidth by the byte convention nnz*(4N + 4 + 4) + sum ndims*R*4
    (indices and value read, residual written, factor rows once)."""
    import torch
    from hiparti_amd import COOTensor
    from hiparti_amd.completion import completion_residuals, cp_completion_loss
    from hiparti_amd.device import DevCOO
    ndims = [500000, 500000, 10000]
    nnz = int(args.cmp_nnz)
    R = 32
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    X0 = COOTensor.synthetic(ndims, nnz, seed=args.seed + 7, nthreads=nthreads)
    inds = X0.inds()
    X0.free()
    gen = torch.Generator(device="cuda").manual_seed(args.seed + 7)
    P = [torch.rand((d, 8), generator=gen, device="cuda", dtype=torch.float64) for d in ndims]
    planted = torch.ones((nnz, 8), dtype=torch.float64, device="cuda")
    for m in range(3):
        planted *= P[m][torch.from_numpy(inds[m].astype("int64")).cuda()]
    vals = planted.sum(dim=1) + 0.01 * torch.randn(nnz, generator=gen, device="cuda", dtype=torch.float64)
    del planted, P
    vals_np = vals.float().cpu().numpy()
    X = COOTensor.from_arrays(ndims, inds, vals_np)
    del vals
    dev = DevCOO(X)
    X.free()
    params = [(0.5 * torch.rand((d, R), generator=gen, device="cuda")).requires_grad_(True) for d in ndims]
    torch.cuda.synchronize()
    log("completion tensor: %d observed entries, ready in %.1f s" % (nnz, time.time() - t0))

    def step():
        for p in params:
            p.grad = None
        loss = cp_completion_loss(dev, params)
        loss.backward()
        return loss

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    det = [p.detach() for p in params]
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in ev:
        a.record()
        completion_residuals(dev, det)
        b.record()
    torch.cuda.synchronize()
    ms_fwd = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    fwd_bytes = nnz * (4 * 3 + 4 + 4) + sum(ndims) * R * 4
    log("completion: %.3f ms per forward+backward, forward %.3f ms, loss %.6g" % (ms_step, ms_fwd, loss.item()))
    res = {"workload": "CP-completion loss, order 3, 500k x 500k x 10k, %d observed entries (uniform), planted "
                       "rank-8 values + N(0, 0.01), R=%d, FP32 factors, FP64 loss" % (nnz, R),
           "ms_per_forward_backward": round(ms_step, 4),
           "ms_per_forward": round(ms_fwd, 4),
           "forward_GBps": round(fwd_bytes / (ms_fwd * 1e-3) / 1e9, 1),
           "forward_hbm_frac": round(fwd_bytes / (ms_fwd * 1e-3) / 8e12, 4),
           "loss": loss.item()}
    if not args.no_errors:
        # The last step's loss and gradients against the oracle's FP64 loss
        # and gradient rows (-MTTKRP of the FP64 residuals).
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.time()
        facs_np = [p.detach().cpu().numpy() for p in params]
        gerr = []
        ref_loss = None
        for mode in range(3):
            rows = sample_rows(inds[mode], ndims[mode], seed=40 + mode)
            ref_loss, ref_rows = O.completion_f64(inds, vals_np, facs_np, mode, rows, nthreads=nthreads)
            got = -params[mode].grad[torch.from_numpy(rows.astype(np.int64)).cuda()].cpu().numpy()
            gerr.append(rel_err(got, ref_rows))
        res["err_vs_fp64"] = {"loss": float("%.3e" % (abs(loss.item() - ref_loss) / ref_loss)),
                              "gradient_rows": [float("%.3e" % e) for e in gerr],
                              "rows": "per mode: the 256 heaviest output rows + 256 random non-empty rows"}
        log("completion errors in %.1f s: %s" % (time.time() - t0, res["err_vs_fp64"]))
    del dev, params
    torch.cuda.empty_cache()
    return {"completion": res}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (MTTKRP would shard by output "
              "block-row ownership); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    import torch
    from hiparti_amd import COOTensor, DenseMatrix, lib  # noqa: F401
    from hiparti_amd.device import DevHiCOO

    assert torch.cuda.is_available(), "no GPU"
    ndims = [1 << 20, 1 << 20, 1 << 16]
    nnz = int(args.nnz)
    R = args.rank
    N = 3
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    setup = {}
    t0 = time.time()
    X = COOTensor.synthetic(ndims, nnz, dist=[1, 1, 1], zipf_s=[1.1] * 3, permute=[1, 1, 1], seed=args.seed,
                            nthreads=nthreads)
    log("generated %d nnz in %.1f s" % (X.nnz(), time.time() - t0))
    inds = vals = rows = None
    if not args.no_errors:
        inds, vals = X.inds(), X.values()
        rows = [sample_rows(inds[m], ndims[m], seed=m) for m in range(N)]
    t0 = time.time()
    H = X.toHiCOO(7, 14, 14, nthreads=nthreads, device=True)
    nb = H.nblocks()
    setup["hicoo_conversion_s"] = round(time.time() - t0, 3)
    log("HiCOO conversion (GPU ordering): nb=%d in %.1f s" % (nb, time.time() - t0))
    t0 = time.time()
    D = DevHiCOO(H)
    torch.cuda.synchronize()
    setup["upload_and_mode_copies_s"] = round(time.time() - t0, 3)
    log("upload %.2f GB in %.1f s" % (D.bytes / 1e9, time.time() - t0))

    facs_np = []
    dfacs = []
    for m, d in enumerate(ndims):
        dm = DenseMatrix(d, R)
        dm.uniform(args.seed * 100 + m)
        f = dm.numpy()
        facs_np.append(f)
        dfacs.append(torch.from_numpy(f).cuda())
    outs = [torch.empty((d, R), dtype=torch.float32, device="cuda") for d in ndims]
    bfacs = [f.to(torch.bfloat16) for f in dfacs]

    def grab():
        return [outs[m][torch.from_numpy(rows[m].astype(np.int64)).cuda()].cpu().numpy() for m in range(N)]

    # The first step also builds each mode's stream form (one-time setup).
    t0 = time.perf_counter()
    for mode in range(N):
        D.mttkrp(dfacs, mode, R=R, out=outs[mode])
    torch.cuda.synchronize()
    setup["first_step_with_stream_build_s"] = round(time.perf_counter() - t0, 3)
    # W - 1 more untimed steps, then exactly K timed steps: ms_per_step is
    # the wall clock of the K steps (synchronised on both sides) over K.
    per_mode, ms_per_step = time_per_mode(D, dfacs, outs, R, N, args.steps, max(args.warmup - 1, 0))
    got32 = grab() if rows is not None else None

    bf16_per_mode = bf16_ms = None
    got16 = None
    if not args.no_bf16:
        t0 = time.perf_counter()
        for mode in range(N):   # builds the BF16 lane shape's stream form when it differs
            D.mttkrp_bf16(bfacs, mode, R=R, out=outs[mode])
        torch.cuda.synchronize()
        setup["first_bf16_step_s"] = round(time.perf_counter() - t0, 3)
        bf16_per_mode, bf16_ms = time_per_mode(D, bfacs, outs, R, N, args.steps, args.warmup, bf16=True)
        got16 = grab() if rows is not None else None


    # Algorithmic bytes per single-mode MTTKRP (see module docstring).
    alg_bytes = hicoo_bytes(nnz, nb, ndims, R)
    avg_ms = float(np.mean(per_mode))
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    attain = attainable_bandwidth()
    log("attainable copy bandwidth %.0f GB/s" % attain)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                "kernel": KERNEL + " (+ output zero fill)",
                "alg_bytes_per_launch": int(alg_bytes),
                "attainable": round(attain, 1),
                "frac_of_attainable": round(achieved / attain, 4)}
    extra = {"per_mode_ms": [round(x, 4) for x in per_mode], "nblocks": nb, "nnz": nnz, "rank": R,
             "setup": setup}
    if bf16_ms is not None:
        bf_bytes = [hicoo_bytes(nnz, nb, ndims, R, mode=m, bf16=True) for m in range(N)]
        extra["bf16_ms_per_step"] = round(bf16_ms, 4)
        extra["bf16_per_mode_ms"] = [round(x, 4) for x in bf16_per_mode]
        extra["bf16_hbm_frac"] = round(float(np.mean([b / (t * 1e-3) for b, t in zip(bf_bytes, bf16_per_mode)]))
                                       / 1e9 / HBM_PEAK_GBS, 4)
        extra["bf16_traffic"] = measured_traffic("bf16_traffic")

    gotcoo = None
    if not args.no_coo:
        coo, gotcoo = coo_leg(X, ndims, dfacs, facs_np, outs, R, args, nthreads, rows=rows)
        extra.update(coo)
    X.free()
    del D, dfacs, outs, bfacs
    torch.cuda.empty_cache()
    if not args.no_precision:
        extra.update(precision_leg(args, nthreads))
    if not args.no_cpd:
        extra.update(cpd_leg(args, nthreads))
    if not args.no_completion:
        extra.update(completion_leg(args))

    # ---- CPU side: accuracy of the timed outputs and the CPU baseline ----
    if rows is not None:
        t0 = time.time()
        e32, e16, ecoo = [], [], []
        rounded = [torch.from_numpy(f).to(torch.bfloat16).float().numpy() for f in facs_np]
        for m in range(N):
            ref = fp64_rows(inds, vals, facs_np, m, rows[m], nthreads)
            e32.append(rel_err(got32[m], ref))
            if gotcoo is not None:
                ecoo.append(rel_err(gotcoo[m], ref))
            if got16 is not None:
                e16.append(rel_err(got16[m], fp64_rows(inds, vals, rounded, m, rows[m], nthreads)))
        extra["err_vs_fp64"] = {"fp32": [float("%.3e" % e) for e in e32],
                                "rows": "per mode: the 256 heaviest output rows + 256 random non-empty rows; "
                                        "BF16 against FP64 on the bf16-rounded factors"}
        if e16:
            extra["err_vs_fp64"]["bf16"] = [float("%.3e" % e) for e in e16]
        if ecoo:
            extra["coo"]["err_vs_fp64"] = [float("%.3e" % e) for e in ecoo]
        log("c1 errors in %.1f s: fp32 %s bf16 %s coo %s" % (time.time() - t0, e32, e16, ecoo))
    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu_ms, kind, z_s, where = cpu_baseline(H, ndims, facs_np, R, args.cpu_sample_nnz, nthreads)
        who = "oracle restatement of ptiOmpMTTKRPHiCOO"
        cpu = {"value": round(cpu_ms, 1), "unit": "ms per step (3-mode HiCOO MTTKRP, extrapolated per nonzero)",
               "cores": nthreads, "kind": kind,
               "sample": "%s (tk=%d, tb=1) on %s = %d nonzeros of the same tensor, all 3 modes, 1 warm-up + 2 "
                         "timed calls per mode" % (who, nthreads, where, z_s)}
        cpu.update(host_facts(nthreads))
        cpu["port_vs_reference"] = PORT_VS_REFERENCE["hicoo"]
        log("cpu baseline %.1f s" % (time.time() - t0))

    line = {
        "metric": "HiCOO MTTKRP ms & % HBM peak (fp32 vs bf16); CPD-ALS s/iter at equal fit",
        "value": round(ms_per_step, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, seed %d)" % args.seed,
        "config": {"workload": "HiCOO MTTKRP all 3 modes, 2^20 x 2^20 x 2^16, %d nnz, Zipf(1.1) relabelled, "
                               "sb=7 sk=14 sc=14, R=%d" % (nnz, R), "nnz": nnz, "rank": R},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
