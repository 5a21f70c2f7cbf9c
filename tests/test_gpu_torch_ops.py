"""torch.ops.gps.cumsum / segmented_cumsum on the GPU against the CPU oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd.torch_ops  # noqa: E402,F401


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def u32_tensor(x: np.ndarray, dtype):
    t = torch.from_numpy(x.view(np.int32)).cuda()
    return t.view(torch.uint32) if dtype == torch.uint32 else t


def u32_numpy(t: torch.Tensor) -> np.ndarray:
    return (t.view(torch.int32) if t.dtype == torch.uint32 else t).cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("dtype", [torch.uint32, torch.int32])
@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("n", [1, 1000, (1 << 21) + 3])
def test_integer_cumsum_is_bit_exact(n, inclusive, dtype):
    x = oracle.gen_u32(n, n, 32)
    y = torch.ops.gps.cumsum(u32_tensor(x, dtype), inclusive=inclusive)
    assert y.dtype == dtype
    assert np.array_equal(u32_numpy(y), oracle.scan_u32(x, inclusive))


def test_float_cumsum_within_one_rounding():
    x = oracle.gen_f32_normal(4, (1 << 20) + 1)
    y = torch.ops.gps.cumsum(torch.from_numpy(x).cuda()).cpu().numpy()
    y64 = oracle.scan_f32_f64(x)
    S = oracle.abs_scan_f32_f64(x)
    assert np.all(np.abs(y - y64) <= 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * S)
    b = oracle.gen_bf16_logmag(5, 70001)
    yb = torch.ops.gps.cumsum(torch.from_numpy(b.view(np.int16)).cuda().view(torch.bfloat16))
    assert yb.dtype == torch.float32
    y64 = oracle.scan_bf16_f64(b)
    assert np.all(np.abs(yb.cpu().numpy() - y64) <= 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * oracle.abs_scan_bf16_f64(b))


def test_segmented_cumsum_integer_and_float():
    off = oracle.gen_lognormal_offsets(9, 300, median=500.0, sigma=1.5, lo=1, hi=1 << 16)
    n = int(off[-1])
    d_off = torch.from_numpy(off.view(np.int64)).cuda()
    x = oracle.gen_u32(3, n, 32)
    y = torch.ops.gps.segmented_cumsum(u32_tensor(x, torch.uint32), d_off, inclusive=False)
    assert np.array_equal(u32_numpy(y), oracle.segmented_scan_u32(x, off, False))
    y = torch.ops.gps.segmented_cumsum(u32_tensor(x, torch.uint32), d_off)  # inclusive by default
    assert np.array_equal(u32_numpy(y), oracle.segmented_scan_u32(x, off, True))
    xf = oracle.gen_f32_normal(6, n)
    yf = torch.ops.gps.segmented_cumsum(torch.from_numpy(xf).cuda(), d_off).cpu().numpy()
    y64 = oracle.segmented_scan_fp_f64(xf, off)
    S = oracle.segmented_scan_fp_f64(xf, off, absolute=True)
    assert np.all(np.abs(yf - y64) <= 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * S)


def test_runs_on_the_current_stream_and_grows_its_handle():
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for n in (1 << 10, 1 << 22, (1 << 23) + 5):  # the cached handle grows past 2^20
            x = oracle.gen_u32(n + 1, n, 32)
            y = torch.ops.gps.cumsum(u32_tensor(x, torch.int32))
            side.synchronize()
            assert np.array_equal(u32_numpy(y), oracle.scan_u32(x, True))


def test_ops_replay_from_a_captured_graph():
    # Warm the (device, stream) handle eagerly at the captured size, capture
    # one integer and one float scan, then replay several times with eager
    # work (including a larger scan that grows the cached handle) in between;
    # every replay must match the oracle.
    import gpuprefixsums_amd.torch_ops as ops
    n = (1 << 21) + 11
    x = oracle.gen_u32(17, n, 32)
    xf = oracle.gen_f32_normal(18, n)
    d_x = u32_tensor(x, torch.int32)
    d_xf = torch.from_numpy(xf).cuda()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        torch.ops.gps.cumsum(d_x)
        torch.ops.gps.cumsum(d_xf)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        y = torch.ops.gps.cumsum(d_x)
        yf = torch.ops.gps.cumsum(d_xf)
    want = oracle.scan_u32(x, True)
    y64 = oracle.scan_f32_f64(xf)
    S = oracle.abs_scan_f32_f64(xf)
    for rep in range(4):
        y.zero_()
        yf.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(u32_numpy(y), want), rep
        assert np.all(np.abs(yf.cpu().numpy() - y64) <= 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * S), rep
        with torch.cuda.stream(s):  # eager work between replays; the second one grows the handle
            m = (1 << 22) + 3 if rep == 1 else 1000
            z = oracle.gen_u32(rep, m, 32)
            assert np.array_equal(u32_numpy(torch.ops.gps.cumsum(u32_tensor(z, torch.int32))),
                                  oracle.scan_u32(z, True))
        s.synchronize()
    assert ops._retired, "the grown handle's predecessor must stay open"


def test_capture_refuses_to_create_a_handle():
    import gpuprefixsums_amd.torch_ops as ops
    n = (1 << 26) + 1  # larger than any handle the other tests create
    x = torch.zeros(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(RuntimeError, match="graph capture"):
        with torch.cuda.graph(g, stream=s):
            torch.ops.gps.cumsum(x)
    # torch reuses pooled streams, so a handle may exist for this stream from
    # an earlier test; it must not have been grown during the capture
    assert all(h.max_n < n for k, h in ops._handles.items() if k[1] == s.cuda_stream)
