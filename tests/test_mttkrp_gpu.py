"""GPU parity of the MTTKRP path (HIP kernels through the C ABI).

Every result is compared with the FP64 oracle (relative Frobenius error
<= 1e-5, the FP32 tolerance) and, where the reference's own FP32 output is in
the fixtures, with that output (<= 1e-5).  Covers the drop-in entry points
ptiMTTKRP / ptiMTTKRPHiCOO on the reference fixtures, orders 2-8, ranks
1-300 (including odd ranks and ranks 2 mod 4 above 64, which COO cuts into
column tiles of 64 lanes x 1 or 2 ranks, and ranks above one 256-rank column
tile), non-default mats_order, empty and single-nonzero
tensors, registered (device-resident) tensors, the device API on torch
tensors, and the BF16-factor form against the FP64 oracle evaluated on the
same bf16-rounded factors.
"""
import numpy as np
import pytest

from conftest import golden_cases, load_golden, rel_fro
from hiparti_amd import COOTensor, DenseMatrix

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


def dense_mats(facs, out_rows, fill=np.nan):
    mats = [DenseMatrix.from_numpy(f) for f in facs]
    out = DenseMatrix(out_rows, facs[0].shape[1])
    out.full_view()[:] = fill
    return mats + [out]


def run_coo(ndims, inds, vals, facs, mode, order=None):
    t = COOTensor.from_arrays(ndims, inds, vals)
    return t.MTTKRP(mode, dense_mats(facs, max(ndims)), order)


def run_hicoo(ndims, inds, vals, facs, mode, sb=7, sk=7, sc=7, order=None):
    t = COOTensor.from_arrays(ndims, inds, vals)
    h = t.toHiCOO(sb, sk, sc)
    return h.MTTKRP(mode, dense_mats(facs, max(ndims)), order)


MTTKRP_CASES = [c for c in golden_cases() if "coo_unif_m0" in load_golden(c)]


@pytest.mark.parametrize("case", MTTKRP_CASES)
def test_dropin_coo_matches_reference_and_fp64(case, oracle):
    g = load_golden(case)
    N = len(g["ndims"])
    R = int(g["R"])
    sets = {"unif": [g["fac%d" % m] for m in range(N)],
            "const": [np.ones((int(d), R), np.float32) for d in g["ndims"]]}
    for kind, facs in sets.items():
        for mode in range(N):
            out = run_coo(g["ndims"], g["in_inds"], g["in_vals"], facs, mode)
            f64 = oracle.mttkrp(g["in_inds"], g["in_vals"], facs, mode, kind="f64")
            assert rel_fro(out, f64) <= TOL, (case, kind, mode)
            assert rel_fro(out, g["coo_%s_m%d" % (kind, mode)]) <= TOL, (case, kind, mode)


@pytest.mark.parametrize("case", MTTKRP_CASES)
def test_dropin_hicoo_matches_reference_and_fp64(case, oracle):
    g = load_golden(case)
    N = len(g["ndims"])
    sb, sk, sc = (int(x) for x in g["params"][0])
    facs = [g["fac%d" % m] for m in range(N)]
    t = COOTensor.from_arrays(g["ndims"], g["in_inds"], g["in_vals"])
    h = t.toHiCOO(sb, sk, sc)
    for mode in range(N):
        out = h.MTTKRP(mode, dense_mats(facs, max(g["ndims"])))
        f64 = oracle.mttkrp(g["in_inds"], g["in_vals"], facs, mode, kind="f64")
        assert rel_fro(out, f64) <= TOL, (case, mode)
        assert rel_fro(out, g["hicoo_unif_m%d" % mode]) <= TOL, (case, mode)


SHAPES = [
    # (ndims, nnz, R, sb, sk)
    ([50, 40], 700, 8, 2, 3),
    ([300, 200, 100], 6000, 1, 3, 5),
    ([300, 200, 100], 6000, 3, 3, 5),
    ([300, 200, 100], 6000, 5, 7, 7),
    ([300, 200, 100], 6000, 32, 2, 4),
    ([300, 200, 100], 6000, 33, 7, 9),
    ([300, 200, 100], 6000, 64, 4, 8),
    ([120, 90, 70], 3000, 100, 3, 5),
    ([120, 90, 70], 3000, 80, 3, 5),
    ([120, 90, 70], 3000, 300, 3, 5),
    # Ranks above 64 whose parity leaves one or two ranks per lane (COO column
    # tiles of 64 x V ranks), and a last HiCOO rank tile narrower than the
    # first tile's vector width.
    ([120, 90, 70], 3000, 65, 3, 5),
    ([120, 90, 70], 3000, 66, 3, 5),
    ([120, 90, 70], 3000, 75, 3, 5),
    ([120, 90, 70], 3000, 127, 3, 5),
    ([120, 90, 70], 3000, 130, 3, 5),
    ([120, 90, 70], 3000, 255, 3, 5),
    ([40, 30, 20, 10], 4000, 16, 2, 3),
    ([30, 20, 15, 10, 8], 5000, 16, 1, 3),
    ([12, 11, 10, 9, 8, 7], 4000, 8, 1, 2),
    ([9, 8, 7, 6, 5, 4, 3], 3000, 12, 1, 1),
    ([6, 6, 5, 5, 4, 4, 3, 3], 3000, 8, 1, 1),
]


@pytest.mark.parametrize("ndims,nnz,R,sb,sk", SHAPES)
def test_random_shapes_against_fp64(ndims, nnz, R, sb, sk, oracle):
    rng = np.random.default_rng(sum(ndims) + R)
    t = COOTensor.synthetic(ndims, nnz, dist=[1] * len(ndims), zipf_s=[1.05] * len(ndims), seed=R)
    inds, vals = t.inds(), t.values()
    N = len(ndims)
    facs = [rng.random((d, R)).astype(np.float32) for d in ndims]
    for mode in range(N):
        f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        assert rel_fro(run_coo(ndims, inds, vals, facs, mode), f64) <= TOL, ("coo", mode)
        assert rel_fro(run_hicoo(ndims, inds, vals, facs, mode, sb, sk, sk), f64) <= TOL, ("hicoo", mode)
    # A different product order changes rounding only.
    order = [N - 1] + list(range(N - 1))[::-1]
    f64 = oracle.mttkrp(inds, vals, facs, N - 1, order=order, kind="f64")
    assert rel_fro(run_hicoo(ndims, inds, vals, facs, N - 1, sb, sk, sk, order=order), f64) <= TOL


def test_empty_and_single_nonzero(oracle):
    ndims = [7, 6, 5]
    facs = [np.random.default_rng(0).random((d, 8)).astype(np.float32) for d in ndims]
    empty = (np.zeros((3, 0), np.uint32), np.zeros(0, np.float32))
    for mode in range(3):
        assert np.all(run_coo(ndims, *empty, facs, mode) == 0)
        assert np.all(run_hicoo(ndims, *empty, facs, mode, 1, 1, 1) == 0)
    one = (np.array([[3], [2], [4]], np.uint32), np.array([2.5], np.float32))
    for mode in range(3):
        f64 = oracle.mttkrp(*one, facs, mode, kind="f64")
        assert rel_fro(run_coo(ndims, *one, facs, mode), f64) <= 1e-7
        assert rel_fro(run_hicoo(ndims, *one, facs, mode, 1, 1, 1), f64) <= 1e-7


def test_output_rows_are_zeroed_and_padding_untouched():
    ndims = [20, 10, 10]
    t = COOTensor.from_arrays(ndims, np.array([[1], [2], [3]], np.uint32), np.ones(1, np.float32))
    facs = [np.ones((d, 5), np.float32) for d in ndims]
    mats = dense_mats(facs, 20, fill=7.0)
    t.MTTKRP(0, mats)
    full = mats[3].full_view()
    expect = np.zeros((20, mats[3].stride()), np.float32)
    expect[1, :5] = 1.0
    np.testing.assert_array_equal(full[:20], expect)


def test_registered_tensors_give_the_same_result(oracle):
    from hiparti_amd import lib
    import ctypes as C
    g = load_golden("tns_3D_12031")
    facs = [g["fac%d" % m] for m in range(3)]
    t = COOTensor.from_arrays(g["ndims"], g["in_inds"], g["in_vals"])
    h = t.toHiCOO(2, 4, 4)
    assert lib().ptiRegisterHiCOO(C.byref(h.address)) == 0
    assert lib().ptiRegisterSparseTensor(C.byref(t.address)) == 0
    try:
        for mode in range(3):
            f64 = oracle.mttkrp(g["in_inds"], g["in_vals"], facs, mode, kind="f64")
            assert rel_fro(h.MTTKRP(mode, dense_mats(facs, 100)), f64) <= TOL
            assert rel_fro(t.MTTKRP(mode, dense_mats(facs, 100)), f64) <= TOL
    finally:
        assert lib().ptiUnregisterHiCOO(C.byref(h.address)) == 0
        assert lib().ptiUnregisterSparseTensor(C.byref(t.address)) == 0


def test_device_api_fp32_and_bf16(oracle):
    import torch
    from hiparti_amd.device import DevCOO, DevHiCOO
    ndims = [5000, 4000, 300]
    R = 32
    t = COOTensor.synthetic(ndims, 200000, dist=[1, 1, 1], zipf_s=[1.1] * 3, permute=[1, 1, 1], seed=3)
    inds, vals = t.inds(), t.values()
    rng = np.random.default_rng(3)
    facs = [rng.random((d, R)).astype(np.float32) for d in ndims]
    dfacs = [torch.from_numpy(f).cuda() for f in facs]
    dc = DevCOO(t)
    h = t.toHiCOO(7, 14, 14)
    dh = DevHiCOO(h)
    assert dh.nnz == 200000
    for mode in range(3):
        f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        out = dh.mttkrp(dfacs, mode)
        torch.cuda.synchronize()
        assert rel_fro(out.cpu().numpy()[:, :R], f64) <= TOL
        out = dc.mttkrp(dfacs, mode)
        torch.cuda.synchronize()
        assert rel_fro(out.cpu().numpy()[:, :R], f64) <= TOL
        # BF16 factors, FP32 products and accumulation: compare with the FP64
        # oracle on the same rounded factors.
        bf = [f.to(torch.bfloat16) for f in dfacs]
        rounded = [b.float().cpu().numpy() for b in bf]
        out = dh.mttkrp_bf16(bf, mode)
        torch.cuda.synchronize()
        f64r = oracle.mttkrp(inds, vals, rounded, mode, kind="f64")
        assert rel_fro(out.cpu().numpy()[:, :R], f64r) <= TOL
        # and its distance to the FP32-factor result is the bf16 rounding (< 1%)
        assert rel_fro(out.cpu().numpy()[:, :R], f64) < 1e-2


@pytest.mark.parametrize("ndims,R", [([3000, 2500, 2000], 16), ([3000, 2500, 2000], 32), ([3000, 2500, 2000], 64),
                                     ([4000, 3000], 32)])
def test_stream_kernel_zipf_ranks_and_orders(ndims, R, oracle):
    """The stream kernel's lane groups of 4, 8 and 16 lanes (ranks 16, 32,
    64: quad_perm, quad_perm plus row shift, and row_newbcast hand-outs) and
    order 2, on Zipf indices with block-rows of thousands of nonzeros (long
    row runs, non-exclusive units).  HiCOO, FP32 and BF16 factors, every
    mode, against the FP64 oracle."""
    import torch
    from hiparti_amd.device import DevHiCOO
    N = len(ndims)
    t = COOTensor.synthetic(ndims, 300000, dist=[1] * N, zipf_s=[1.1] * N, permute=[1] * N, seed=11)
    inds, vals = t.inds(), t.values()
    rng = np.random.default_rng(12)
    facs = [rng.random((d, R)).astype(np.float32) for d in ndims]
    dfacs = [torch.from_numpy(f).cuda() for f in facs]
    dh = DevHiCOO(t.toHiCOO(7, 14, 14))
    bf = [f.to(torch.bfloat16) for f in dfacs]
    rounded = [b.float().cpu().numpy() for b in bf]
    for mode in range(N):
        f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        out = dh.mttkrp(dfacs, mode)
        torch.cuda.synchronize()
        assert rel_fro(out.cpu().numpy()[:, :R], f64) <= TOL, mode
        if R == 32:
            out = dh.mttkrp_bf16(bf, mode)
            torch.cuda.synchronize()
            assert rel_fro(out.cpu().numpy()[:, :R], oracle.mttkrp(inds, vals, rounded, mode, kind="f64")) <= TOL, mode


@pytest.mark.parametrize("ndims,R", [([3000, 2500, 2000], 32), ([3000, 2500, 2000], 40), ([3000, 2500, 2000], 64),
                                     ([3000, 2500, 2000], 96), ([3000, 2500, 2000], 128), ([400, 300, 200, 50], 64),
                                     ([300, 200, 100, 40, 20], 32), ([4000, 3000], 48)])
def test_bf16_eight_ranks_per_lane(ndims, R, oracle):
    """BF16 factors at ranks that are multiples of 8 and at least 32 gather
    eight ranks (16 bytes) per lane on the stream kernel, in lane groups of 4
    or 8 lanes with a stream form of their own; ranks above 64 cut into column
    tiles that share the first tile's lane groups.  Against the FP64 oracle on
    the same bf16-rounded factors, every mode, with FP32 calls interleaved (the
    FP32 and BF16 stream forms of a mode are cached side by side)."""
    import torch
    from hiparti_amd.device import DevHiCOO
    N = len(ndims)
    t = COOTensor.synthetic(ndims, 200000, dist=[1] * N, zipf_s=[1.1] * N, permute=[1] * N, seed=R + N)
    inds, vals = t.inds(), t.values()
    rng = np.random.default_rng(R)
    facs = [rng.random((d, R)).astype(np.float32) for d in ndims]
    dfacs = [torch.from_numpy(f).cuda() for f in facs]
    dh = DevHiCOO(t.toHiCOO(7, 14, 14))
    bf = [f.to(torch.bfloat16) for f in dfacs]
    rounded = [b.float().cpu().numpy() for b in bf]
    for rep in range(2):
        for mode in range(N):
            out = dh.mttkrp_bf16(bf, mode)
            torch.cuda.synchronize()
            ref = oracle.mttkrp(inds, vals, rounded, mode, kind="f64")
            assert rel_fro(out.cpu().numpy()[:, :R], ref) <= TOL, (rep, mode)
            out = dh.mttkrp(dfacs, mode)
            torch.cuda.synchronize()
            assert rel_fro(out.cpu().numpy()[:, :R], oracle.mttkrp(inds, vals, facs, mode, kind="f64")) <= TOL


def test_large_zipf_properties(oracle):
    """Size-independent checks on a c1-shaped tensor (Zipf 1.1, relabelled
    modes, 2^20 x 2^20 x 2^16, sb=7 sk=14) at 4M nonzeros: all-ones factors
    give segmented sums; HiCOO and COO paths agree with the FP64 oracle."""
    ndims = [1 << 20, 1 << 20, 1 << 16]
    t = COOTensor.synthetic(ndims, 4_000_000, dist=[1, 1, 1], zipf_s=[1.1] * 3, permute=[1, 1, 1], seed=42)
    inds, vals = t.inds(), t.values()
    R = 32
    h = t.toHiCOO(7, 14, 14)
    ones = [np.ones((d, R), np.float32) for d in ndims]
    rng = np.random.default_rng(9)
    facs = [rng.random((d, R)).astype(np.float32) for d in ndims]
    for mode in range(3):
        seg = np.bincount(inds[mode], weights=vals.astype(np.float64), minlength=ndims[mode])
        out = h.MTTKRP(mode, dense_mats(ones, max(ndims)))
        assert rel_fro(out, np.repeat(seg[:, None], R, axis=1)) <= TOL
        f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        assert rel_fro(h.MTTKRP(mode, dense_mats(facs, max(ndims))), f64) <= TOL
        assert rel_fro(t.MTTKRP(mode, dense_mats(facs, max(ndims))), f64) <= TOL


def test_device_upload_rejects_malformed_hicoo():
    """The device layout build validates the arrays it uses as memory and LDS
    indices: element offsets < 2^sb_bits, coordinates inside their modes,
    non-empty blocks."""
    import ctypes as C
    from hiparti_amd import PtiError
    from hiparti_amd import _ffi as F
    from hiparti_amd.device import DevHiCOO
    rng = np.random.default_rng(17)
    ndims = [60, 50, 40]
    inds = np.stack([rng.integers(0, d, 3000) for d in ndims]).astype(np.uint32)
    vals = rng.random(3000).astype(np.float32)
    for corrupt in ("eind_too_big", "coord_outside", "empty_block"):
        t = COOTensor.from_arrays(ndims, inds, vals)
        h = t.toHiCOO(3, 4, 4)
        a = h.address
        nnz = int(a.nnz)
        if corrupt == "eind_too_big":
            np.ctypeslib.as_array(a.einds[1].data, shape=(nnz,))[nnz // 2] = 8      # sb_bits = 3
        elif corrupt == "coord_outside":
            nb = int(a.bptr.len) - 1
            np.ctypeslib.as_array(a.binds[2].data, shape=(nb,))[nb // 3] = 5      # (5 << 3) >= 40
        else:
            bp = np.ctypeslib.as_array(a.bptr.data, shape=(int(a.bptr.len),))
            bp[2] = bp[1]
        with pytest.raises(PtiError) as e:
            DevHiCOO(h)
        assert e.value.code == F.PTIERR_VALUE_ERROR, corrupt


def test_order3_factors_of_4gib_take_the_tile_kernel():
    """Order-3 MTTKRPs whose gathered factor matrices span 4 GiB or more run
    the tile kernel (the stream kernel gathers with 32-bit byte offsets); mode
    0 here gathers only small factors and runs the stream kernel.  Both
    against the FP64 definition on the nonzeros' factor rows."""
    import torch
    from hiparti_amd.device import DevHiCOO
    ndims = [1 << 25, 300, 200]     # mode-0 factor: 2^25 x 32 FP32 = 4 GiB
    R = 32
    rng = np.random.default_rng(23)
    nnz = 20000
    inds = np.stack([rng.integers(0, d, nnz) for d in ndims]).astype(np.uint32)
    t = COOTensor.from_arrays(ndims, inds, rng.random(nnz).astype(np.float32))
    dh = DevHiCOO(t.toHiCOO(7, 10, 10))
    inds, vals = t.inds().astype(np.int64), t.values().astype(np.float64)
    gen = torch.Generator(device="cuda").manual_seed(5)
    dfacs = [torch.rand((d, R), generator=gen, device="cuda") for d in ndims]
    rows = [dfacs[m][torch.from_numpy(inds[m]).cuda()].double().cpu().numpy() for m in range(3)]
    for mode in range(3):
        a, b = [m for m in range(3) if m != mode]
        want = np.zeros((ndims[mode], R))
        np.add.at(want, inds[mode], vals[:, None] * rows[a] * rows[b])
        out = dh.mttkrp(dfacs, mode)
        torch.cuda.synchronize()
        touched = np.unique(inds[mode])
        got = out[torch.from_numpy(touched).cuda()].cpu().numpy()
        assert rel_fro(got, want[touched]) <= TOL, mode
        assert float(out.abs().sum().item()) == pytest.approx(float(np.abs(want).sum()), rel=1e-5)
        del out


def test_stream_kernel_with_coordinates_beyond_24_bits():
    """A non-target mode with more than 2^24 rows sends the stream kernel's
    row offsets through the 32-bit multiply (the 24-bit multiply-add covers
    smaller modes): modes 1 and 2 gather rows of mode 0 at and above 2^24,
    HiCOO and COO, FP32, against the FP64 definition."""
    import torch
    from hiparti_amd.device import DevCOO, DevHiCOO
    ndims = [(1 << 24) + 40, 300, 200]   # mode-0 factor: 2^24 x 16 FP32 = 1 GiB
    R = 16
    rng = np.random.default_rng(29)
    nnz = 30000
    inds = np.stack([rng.integers(0, d, nnz) for d in ndims]).astype(np.uint32)
    inds[0, : nnz // 3] = rng.integers((1 << 24) - 8, ndims[0], nnz // 3)
    keys = np.unique(inds[0].astype(np.uint64) * 60000 + inds[1].astype(np.uint64) * 200 + inds[2], return_index=True)[1]
    inds = inds[:, keys]
    t = COOTensor.from_arrays(ndims, inds, rng.random(inds.shape[1]).astype(np.float32))
    dc = DevCOO(t)
    dh = DevHiCOO(t.toHiCOO(7, 10, 10))
    inds, vals = t.inds().astype(np.int64), t.values().astype(np.float64)
    gen = torch.Generator(device="cuda").manual_seed(6)
    dfacs = [torch.rand((d, R), generator=gen, device="cuda") for d in ndims]
    rows = [dfacs[m][torch.from_numpy(inds[m]).cuda()].double().cpu().numpy() for m in range(3)]
    for mode in (1, 2):
        a, b = [m for m in range(3) if m != mode]
        want = np.zeros((ndims[mode], R))
        np.add.at(want, inds[mode], vals[:, None] * rows[a] * rows[b])
        for name, d in (("hicoo", dh), ("coo", dc)):
            got = d.mttkrp(dfacs, mode).cpu().numpy()
            assert rel_fro(got, want) <= TOL, (name, mode)


@pytest.mark.parametrize("ndims", [[(1 << 24) + 40, 300, 200, 50], [(1 << 24) + 40, 300, 200, 50, 20],
                                   [3000, 300, 200, 50, 70000], [3000, 300, 200, 50, 1000]])
def test_packed_and_two_plane_stream_forms(ndims):
    """Order-4 and order-5 stream forms pack each entry into one 16-byte word
    (the output row in the first coordinate's high byte; for order 5 the last
    gathered coordinate split over the second and third words' high bytes)
    only where every gathered coordinate is below 2^24 and, for order 5, the
    last gathered mode has at most 2^16 rows; other forms keep two planes.
    Here the modes that gather the (2^24 + 40)-row mode 0 (some coordinates
    at and above 2^24) and, in the third tensor, modes 0-3 (whose last
    gathered mode has 70000 rows) take two planes; the other modes pack, in
    the fourth tensor with last gathered coordinates up to 999 (both bytes
    used), some with the short trailing factor resident in LDS.  HiCOO FP32
    and BF16 against the FP64 definition."""
    import torch
    from hiparti_amd.device import DevHiCOO
    N = len(ndims)
    R = 16
    rng = np.random.default_rng(31)
    nnz = 40000
    inds = np.stack([rng.integers(0, d, nnz) for d in ndims]).astype(np.uint32)
    if ndims[0] > (1 << 24):
        inds[0, : nnz // 3] = rng.integers((1 << 24) - 8, ndims[0], nnz // 3)
    key = np.zeros(nnz, dtype=np.uint64)
    for m in range(N):
        key = key * np.uint64(ndims[m] if m else 1) + inds[m].astype(np.uint64)
    inds = inds[:, np.unique(key, return_index=True)[1]]
    t = COOTensor.from_arrays(ndims, inds, rng.random(inds.shape[1]).astype(np.float32))
    dh = DevHiCOO(t.toHiCOO(7, 10, 10))
    inds, vals = t.inds().astype(np.int64), t.values().astype(np.float64)
    gen = torch.Generator(device="cuda").manual_seed(8)
    dfacs = [torch.rand((d, R), generator=gen, device="cuda") for d in ndims]
    bf = [f.to(torch.bfloat16) for f in dfacs]
    for facs, name in ((dfacs, "fp32"), (bf, "bf16")):
        rows = [facs[m][torch.from_numpy(inds[m]).cuda()].double().cpu().numpy() for m in range(N)]
        for mode in range(N):
            want = np.zeros((ndims[mode], R))
            prod = vals[:, None].copy()
            for m in range(N):
                if m != mode:
                    prod = prod * rows[m]
            np.add.at(want, inds[mode], prod)
            got = (dh.mttkrp(facs, mode) if name == "fp32" else dh.mttkrp_bf16(facs, mode)).cpu().numpy()[:, :R]
            assert rel_fro(got, want) <= TOL, (N, name, mode)


def test_stream_paths_on_extreme_row_distributions(oracle):
    """The stream kernels' run sums, chunk padding and unit write-out on
    extreme shapes, HiCOO and COO against the FP64 oracle: every nonzero in
    one output row (one run across many chunks and non-exclusive units),
    block-rows of exactly 256 * k nonzeros (no padding), and ranks 4 and 60
    (idle lanes)."""
    rng = np.random.default_rng(41)
    ndims = [300, 2000, 1500]
    # All 60000 nonzeros share output row 7 of mode 0.
    nnz = 60000
    flat = rng.choice(ndims[1] * ndims[2], nnz, replace=False)
    one_row = np.stack([np.full(nnz, 7), flat // ndims[2], flat % ndims[2]]).astype(np.uint32)
    # Each of 8 block-rows (sb = 7) holds exactly 512 nonzeros.
    blocks = []
    for br in range(8):
        f = rng.choice(128 * ndims[1] * 10, 512, replace=False)
        blocks.append(np.stack([br * 128 + f % 128, (f // 128) % ndims[1], f // (128 * ndims[1])]))
    full = np.concatenate(blocks, axis=1).astype(np.uint32)
    for name, nd, inds in (("one row", ndims, one_row), ("full chunks", [1024, 2000, 10], full)):
        vals = rng.random(inds.shape[1]).astype(np.float32)
        for R in (4, 60, 32):
            facs = [rng.random((d, R)).astype(np.float32) for d in nd]
            for mode in range(3):
                f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
                assert rel_fro(run_hicoo(nd, inds, vals, facs, mode, 7, 10, 10), f64) <= TOL, (name, R, mode)
                assert rel_fro(run_coo(nd, inds, vals, facs, mode), f64) <= TOL, (name, R, mode, "coo")
