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
fill), against 8 TB/s; the BF16-factor form's time; and the oracle's
OpenMP HiCOO MTTKRP (the reference's superblock-parallel scheme) timed on a
bounded sample of the same tensor.  roofline.traffic is the HBM bytes per
launch from the committed rocprofv3 counter summary of this kernel (null if
that summary was taken on another kernel version).

The "coo" object is ptiMTTKRP's COO kernel on the same tensor and factors
(ms per step, its roofline with the COO byte convention, and the reference's
ptiOmpMTTKRP timed on a sample of the same tensor).

The "cpd" object is the CPD-ALS leg: an order-5 tensor 1M x 500k x 100k x
1k x 24 with 150M nonzeros, rank 16, 20 iterations on the device from one
seeded set of factors, all-FP32 versus the BF16-then-FP32 schedule (s/iter
and final fit of each).

    python bench.py [--steps K] [--warmup W] [--nnz NNZ] [--no-cpu] [--no-cpd] [--no-coo]
"""
import argparse
import ctypes as C
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


KERNEL = "hicoo_stream_kernel<3, 4, 8, float, 8>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hicoo_stream_kernel_zipf100M_R32_counters.md")


def measured_traffic():
    """HBM bytes per launch of the FP32 HiCOO kernel on this workload, from
    the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE,
    calibrated on a plain copy), when it was taken on the current kernel."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    kern = re.search(r"^kernel: (.*)$", text, re.M)
    val = re.search(r"^traffic_bytes_per_launch: (\d+)$", text, re.M)
    if not kern or not val or kern.group(1).strip() != KERNEL:
        return None
    return int(val.group(1))


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libhiparti_ref.so")


def _sample_prefix(h, sample_nnz):
    """Smallest prefix of superblocks holding >= sample_nnz nonzeros:
    (k_end superblocks, b_end blocks, z_end nonzeros)."""
    a = h.address
    nk_total = int(a.kptr.len) - 1
    kptr = np.ctypeslib.as_array(a.kptr.data, shape=(nk_total + 1,))
    bptr_all = np.ctypeslib.as_array(a.bptr.data, shape=(int(a.bptr.len),))
    k_end = int(np.searchsorted(bptr_all[kptr], sample_nnz))
    k_end = max(1, min(k_end, nk_total))
    b_end = int(kptr[k_end])
    return k_end, b_end, int(bptr_all[b_end])


def _time_modes(run, N):
    """One warm-up and two timed calls per mode; seconds per three-mode pass."""
    total = 0.0
    for mode in range(N):
        run(mode)
        t0 = time.perf_counter()
        for _ in range(2):
            run(mode)
        total += (time.perf_counter() - t0) / 2
    return total


def cpu_baseline(h, ndims, facs_np, R, sample_nnz, nthreads):
    """The CPU baseline on the leading superblocks holding about sample_nnz
    nonzeros of the same HiCOO tensor, one warm-up and two timed calls per
    mode, extrapolated per nonzero to the whole tensor (ms per three-mode
    step).  It is the reference's own ptiOmpMTTKRPHiCOO with tk = host
    threads and tb = 1 (superblock-parallel with per-element atomics,
    src/sptensor/hicoo/mttkrp_omp.c:184-204,400-467) when its library has been
    built from the reference sources (oracle/_ref), else the oracle's
    restatement of that kernel.  Returns (ms, kind, z_end, k_end)."""
    from hiparti_amd import _ffi as F
    a = h.address
    N = len(ndims)
    k_end, b_end, z_end = _sample_prefix(h, sample_nnz)
    if os.path.exists(REF_LIB):
        L = C.CDLL(REF_LIB)
        L.ptiOmpMTTKRPHiCOO.argtypes = [C.POINTER(F.ptiSparseTensorHiCOO), C.POINTER(C.POINTER(F.ptiMatrix)),
                                        C.POINTER(F.ptiIndex), F.ptiIndex, C.c_int, C.c_int]
        L.ptiOmpMTTKRPHiCOO.restype = C.c_int
        view = F.ptiSparseTensorHiCOO.from_buffer_copy(a)   # same arrays, first k_end superblocks
        view.kptr.len = k_end + 1
        view.nnz = z_end
        mats = []
        for f in facs_np:
            m = F.ptiMatrix(f.shape[0], f.shape[1], f.shape[0], f.shape[1], f.ctypes.data_as(C.POINTER(F.ptiValue)))
            mats.append(m)
        out = np.zeros((max(ndims), R), np.float32)
        mats.append(F.ptiMatrix(max(ndims), R, max(ndims), R, out.ctypes.data_as(C.POINTER(F.ptiValue))))
        ptrs = (C.POINTER(F.ptiMatrix) * (N + 1))(*[C.pointer(m) for m in mats])

        def run(mode):
            order = (F.ptiIndex * N)(*([mode] + [(mode + i) % N for i in range(1, N)]))
            mats[N].nrows = ndims[mode]
            assert L.ptiOmpMTTKRPHiCOO(C.byref(view), ptrs, order, mode, int(nthreads), 1) == 0
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        kptr = np.ctypeslib.as_array(a.kptr.data, shape=(int(a.kptr.len),))
        sub = {
            "kptr": kptr[:k_end + 1].copy(),
            "bptr": np.ctypeslib.as_array(a.bptr.data, shape=(b_end + 1,)).copy(),
            "binds": np.stack([np.ctypeslib.as_array(a.binds[m].data, shape=(b_end,)) for m in range(N)]),
            "einds": np.stack([np.ctypeslib.as_array(a.einds[m].data, shape=(z_end,)) for m in range(N)]),
            "values": np.ctypeslib.as_array(a.values.data, shape=(z_end,)).copy(),
        }

        def run(mode):
            O.mttkrp_hicoo(sub, int(a.sb_bits), facs_np, mode, kind="omp", nthreads=nthreads)
        kind = "port"
    total = _time_modes(run, N)
    return total / z_end * int(a.nnz) * 1e3, kind, z_end, k_end


def coo_leg(X, ndims, dfacs, facs_np, outs, R, args, nthreads):
    """ptiMTTKRP on the device (COO, per-mode copies sorted by the output
    index, cut into 128-row blocks and run by the stream kernel) on the same
    tensor and factors: ms per three-mode step with HIP events,
    its roofline with the reference's COO convention B = nnz*(N*4 + 4) +
    sum_m ndims[m]*R*4 (benchmark/tensor/mttkrp.c:304-307), and the CPU
    baseline: the reference's own ptiOmpMTTKRP (src/sptensor/mttkrp_omp.c:
    125-205, nonzero-parallel with atomics) on the first cpu_sample_nnz
    nonzeros, extrapolated per nonzero (the oracle's restatement where the
    reference library is absent)."""
    import torch
    from hiparti_amd import _ffi as F
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
    del Dc
    torch.cuda.empty_cache()
    alg = nnz * (N * 4 + 4) + sum(d * R * 4 for d in ndims)
    achieved = alg / (np.mean(per_mode) * 1e-3) / 1e9
    res = {"ms_per_step": round(ms, 4), "per_mode_ms": [round(x, 4) for x in per_mode],
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "hicoo_stream_kernel<3, 4, 8, float, 8> on the COO copy's 128-row blocks "
                                  "(+ output zero fill)",
                        "alg_bytes_per_launch": int(alg)}}
    if not args.no_cpu:
        t0 = time.time()
        z = int(min(args.cpu_sample_nnz, nnz))
        if os.path.exists(REF_LIB):
            L = C.CDLL(REF_LIB)
            L.ptiOmpMTTKRP.argtypes = [C.POINTER(F.ptiSparseTensor), C.POINTER(C.POINTER(F.ptiMatrix)),
                                       C.POINTER(F.ptiIndex), F.ptiIndex, C.c_int]
            L.ptiOmpMTTKRP.restype = C.c_int
            view = F.ptiSparseTensor.from_buffer_copy(X.address)   # the first z nonzeros
            view.nnz = z
            mats = [F.ptiMatrix(f.shape[0], f.shape[1], f.shape[0], f.shape[1],
                                f.ctypes.data_as(C.POINTER(F.ptiValue))) for f in facs_np]
            out = np.zeros((max(ndims), R), np.float32)
            mats.append(F.ptiMatrix(max(ndims), R, max(ndims), R, out.ctypes.data_as(C.POINTER(F.ptiValue))))
            ptrs = (C.POINTER(F.ptiMatrix) * (N + 1))(*[C.pointer(m) for m in mats])

            def run(mode):
                order = (F.ptiIndex * N)(*([mode] + [(mode + i) % N for i in range(1, N)]))
                mats[N].nrows = ndims[mode]
                assert L.ptiOmpMTTKRP(C.byref(view), ptrs, order, mode, int(nthreads)) == 0
            kind, who = "reference", "reference ptiOmpMTTKRP"
        else:
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
        log("coo cpu baseline %.1f s" % (time.time() - t0))
    return {"coo": res}


def cpd_leg(args):
    """CP-ALS on the device: order-5 tensor 1M x 500k x 100k x 1k x 24,
    Zipf(1.1) indices on modes 0-2 and uniform on modes 3-4, values U(0,1),
    HiCOO sb=7 sk=14 sc=14, rank 16, a fixed number of iterations from the
    same seeded U(0,1) factors: all-FP32 MTTKRPs, then the schedule whose
    first iterations gather BF16 factor copies.  Returns s/iter and the final
    fits of both runs."""
    import torch
    from hiparti_amd import COOTensor
    from hiparti_amd import cpd as CP
    from hiparti_amd.device import DevHiCOO
    ndims = [1000000, 500000, 100000, 1000, 24]
    nnz = int(args.cpd_nnz)
    R = args.cpd_rank
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    X = COOTensor.synthetic(ndims, nnz, dist=[1, 1, 1, 0, 0], zipf_s=[1.1, 1.1, 1.1, 0, 0],
                            permute=[0, 0, 0, 0, 0], seed=args.seed + 4, nthreads=nthreads)
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
    for name, bf in (("fp32", 0), ("bf16_fp32", args.cpd_bf16_iters)):
        facs = [f.clone() for f in init]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = CP.cp_als(D, R, niters=args.cpd_iters, tol=0.0, factors=facs, bf16_iters=bf)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        runs[name] = (sec / max(res.iters, 1), res.fit, res.iters)
        log("cpd %s: %.4f s/iter, fit %.6f after %d iterations" % (name, runs[name][0], res.fit, res.iters))
    del D
    torch.cuda.empty_cache()
    return {"cpd": {"workload": "CP-ALS order-5 1M x 500k x 100k x 1k x 24, %d nnz, Zipf(1.1) modes 0-2, "
                                "uniform modes 3-4, R=%d, %d iterations, HiCOO sb=7 sk=14" % (nnz, R, args.cpd_iters),
                    "nblocks": nb,
                    "fp32_s_per_iter": round(runs["fp32"][0], 5), "fp32_fit": runs["fp32"][1],
                    "bf16_iters": args.cpd_bf16_iters,
                    "bf16_fp32_s_per_iter": round(runs["bf16_fp32"][0], 5), "bf16_fp32_fit": runs["bf16_fp32"][1],
                    "fit_delta": abs(runs["fp32"][1] - runs["bf16_fp32"][1])}}


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
    X = COOTensor.from_arrays(ndims, inds, vals.float().cpu().numpy())
    del vals, inds
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
    del dev, params
    torch.cuda.empty_cache()
    return {"completion": {
        "workload": "CP-completion loss, order 3, 500k x 500k x 10k, %d observed entries (uniform), planted rank-8 "
                    "values + N(0, 0.01), R=%d, FP32 factors, FP64 loss" % (nnz, R),
        "ms_per_forward_backward": round(ms_step, 4),
        "ms_per_forward": round(ms_fwd, 4),
        "forward_GBps": round(fwd_bytes / (ms_fwd * 1e-3) / 1e9, 1),
        "forward_hbm_frac": round(fwd_bytes / (ms_fwd * 1e-3) / 8e12, 4),
        "loss": loss.item()}}


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
    from hiparti_amd import COOTensor, DenseMatrix, lib
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

    def step(bf16=False, events=None):
        for mode in range(N):
            if events is not None:
                events[mode][0].record()
            if bf16:
                D.mttkrp_bf16(bfacs, mode, R=R, out=outs[mode])
            else:
                D.mttkrp(dfacs, mode, R=R, out=outs[mode])
            if events is not None:
                events[mode][1].record()

    # The first step also builds each mode's stream form (one-time setup).
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    setup["first_step_with_stream_build_s"] = round(time.perf_counter() - t0, 3)
    for _ in range(max(args.warmup - 1, 0)):
        step()
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(N)]
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events=ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_mode = [float(np.mean([ev[k][m][0].elapsed_time(ev[k][m][1]) for k in range(args.steps)]))
                for m in range(N)]
    ms_per_step = wall * 1e3 / args.steps

    bf16_ms = None
    bf16_per_mode = None
    if not args.no_bf16:
        for _ in range(args.warmup):
            step(bf16=True)
        torch.cuda.synchronize()
        evb = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(N)]
               for _ in range(args.steps)]
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(bf16=True, events=evb[k])
        torch.cuda.synchronize()
        bf16_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        bf16_per_mode = [float(np.mean([evb[k][m][0].elapsed_time(evb[k][m][1]) for k in range(args.steps)]))
                         for m in range(N)]

    # Algorithmic bytes per single-mode MTTKRP (see module docstring).
    fac_bytes = sum(d * R * 4 for d in ndims)
    alg_bytes = nnz * (N * 1 + 4) + nb * (N * 4 + 8) + fac_bytes
    avg_ms = float(np.mean(per_mode))
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                "kernel": KERNEL + " (+ output zero fill)",
                "alg_bytes_per_launch": int(alg_bytes)}
    extra = {"per_mode_ms": [round(x, 4) for x in per_mode], "nblocks": nb, "nnz": nnz, "rank": R,
             "setup": setup}
    if bf16_ms is not None:
        bf_bytes = nnz * (N * 1 + 4) + nb * (N * 4 + 8) + sum(d * R * 2 for d in ndims)
        extra["bf16_ms_per_step"] = round(bf16_ms, 4)
        extra["bf16_per_mode_ms"] = [round(x, 4) for x in bf16_per_mode]
        extra["bf16_hbm_frac"] = round(bf_bytes / (np.mean(bf16_per_mode) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)

    if not args.no_coo:
        extra.update(coo_leg(X, ndims, dfacs, facs_np, outs, R, args, nthreads))
    X.free()
    cpu = None
    cpu_args = (H, ndims, facs_np, R)
    if not args.no_cpd:
        del D, dfacs, outs, bfacs
        torch.cuda.empty_cache()
        extra.update(cpd_leg(args))
    if not args.no_completion:
        extra.update(completion_leg(args))
    if not args.no_cpu:
        t0 = time.time()
        cpu_ms, kind, z_s, k_s = cpu_baseline(*cpu_args, args.cpu_sample_nnz, nthreads)
        who = "reference ptiOmpMTTKRPHiCOO" if kind == "reference" else "oracle restatement of ptiOmpMTTKRPHiCOO"
        cpu = {"value": round(cpu_ms, 1), "unit": "ms per step (3-mode HiCOO MTTKRP, extrapolated per nonzero)",
               "cores": nthreads, "kind": kind,
               "sample": "%s (tk=%d, tb=1) on the first %d superblocks = %d nonzeros of the same tensor, all 3 "
                         "modes, 1 warm-up + 2 timed calls per mode" % (who, nthreads, k_s, z_s)}
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
