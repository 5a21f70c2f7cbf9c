"""Use patterns beyond one scan at a time: a plain scan captured in a HIP
graph (through torch.cuda.graph) and replayed with eager work in between,
two handles scanning concurrently on two streams, and the reference-style
test program (the reference's main(), GPUPrefixSumsCUDA.cu:14-38)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


@pytest.mark.parametrize("schedule", [gps.SCHEDULE_PERSISTENT, gps.SCHEDULE_TILE])
def test_plain_scan_graph_replays_match_the_oracle(schedule):
    # One launch per scan and no memset inside the call, so the scan can be
    # captured; the epoch in the ticket word keeps replays and the eager calls
    # between them apart. Every replay is checked.
    n = (1 << 22) + 3
    x = oracle.gen_u32(41, n, 32)
    want_incl, want_excl = oracle.scan_u32(x, True), oracle.scan_u32(x, False)
    d_in, d_out, d_eager = to_dev(x), sentinel_u32(n), sentinel_u32(n)
    with gps.ScanHandle(gps.ALGO_CSDL, n, schedule=schedule) as h:
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):  # warm-up on the capturing stream (one-time occupancy queries)
                h.inclusive(d_in, d_out, n, stream=side)
        torch.cuda.current_stream().wait_stream(side)
        sync()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            h.inclusive(d_in, d_out, n, stream=torch.cuda.current_stream())
        for rep in range(4):
            d_out.fill_(0)
            g.replay()
            sync()
            got = from_dev_u32(d_out, n)
            assert np.array_equal(got, want_incl), f"replay {rep}: {first_mismatch(got, want_incl)}"
            h.exclusive(d_in, d_eager, n)  # eager work on the same handle between replays
            sync()
            got = from_dev_u32(d_eager, n)
            assert np.array_equal(got, want_excl), f"eager after replay {rep}: {first_mismatch(got, want_excl)}"
        assert h.device_error() == 0


def test_two_handles_scan_concurrently_on_two_streams():
    # Two persistent grids, each sized to the whole device, run at once: the
    # second one's workgroups wait for resources. Tickets go only to running
    # workgroups, so neither scan can wait on a workgroup that never started.
    n = 1 << 24
    xs = [oracle.gen_u32(50 + i, n, 32) for i in range(2)]
    wants = [oracle.scan_u32(x, i == 0) for i, x in enumerate(xs)]
    d_in = [to_dev(x) for x in xs]
    d_out = [sentinel_u32(n) for _ in xs]
    streams = [torch.cuda.Stream() for _ in xs]
    for s in streams:
        s.wait_stream(torch.cuda.current_stream())
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h0, gps.ScanHandle(gps.ALGO_CSDL, n) as h1:
        for _ in range(6):
            h0.inclusive(d_in[0], d_out[0], n, stream=streams[0])
            h1.exclusive(d_in[1], d_out[1], n, stream=streams[1])
        sync()
        for i in range(2):
            got = from_dev_u32(d_out[i], n)
            assert np.array_equal(got, wants[i]), f"handle {i}: {first_mismatch(got, wants[i])}"
        assert h0.device_error() == 0 and h1.device_error() == 0


def test_reference_style_main_runs_and_passes():
    exe = ROOT / "gpuprefixsums_amd" / "gps_main"
    if not exe.exists():
        pytest.fail("gps_main is not built (python -m gpuprefixsums_amd.build)")
    res = subprocess.run([str(exe), "22", "3"], capture_output=True, text=True, timeout=240)
    out = res.stdout
    assert res.returncode == 0, out[-2000:] + res.stderr[-2000:]
    assert out.count("All tests passed.") == 6, out[-2000:]  # CSDL, RTS, CSDLDF x inclusive/exclusive
    assert "rocPRIM inclusive scan batch timing" in out
