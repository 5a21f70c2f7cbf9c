"""The memory-ceiling probes (libgps_ceiling.so) move exactly the bytes they
claim: each copy reproduces its input and writes nothing past it, the write
stream and fill cover their range and stop there, and the ticketed copies
re-arm their ticket word on every launch. The benchmark reads the scan's
roofline against these probes, so a probe that skipped work would flatter
the scan.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from gpuprefixsums_amd import ceiling  # noqa: E402
from tests.gpu_util import from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402

SENTINEL = 0xDEADBEEF


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def random_words(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("kind", ceiling.COPY_KINDS, ids=lambda k: f"kind{k}")
@pytest.mark.parametrize("tiles", [1, 3, 1000])
def test_copy_probes_copy_exactly(kind, tiles):
    nbytes = tiles * max(ceiling.granule(kind), 128 * 1024)  # whole tiles of every copy kind
    n = nbytes // 4
    x = random_words(n, seed=tiles + 17 * kind)
    src = to_dev(x)
    dst = sentinel_u32(n + 256)
    scratch = torch.zeros(max(ceiling.scratch_bytes(kind, nbytes), 4) // 4, dtype=torch.int32, device="cuda")
    assert scratch.data_ptr() % 256 == 0
    for _ in range(3):  # repeated launches re-use the re-armed ticket
        ceiling.run(kind, src, dst, nbytes, scratch)
    sync()
    out = from_dev_u32(dst)
    assert np.array_equal(out[:n], x), f"{ceiling.name(kind)}: copy differs"
    assert np.all(out[n:] == SENTINEL), f"{ceiling.name(kind)}: wrote past the range"
    if kind in ceiling.TICKETED:
        words = from_dev_u32(scratch)
        assert words[0] == 0, "ticket word not re-armed"
        if kind in (ceiling.TICKET_1024X4_STATUS, ceiling.TICKET_1024X8_STATUS):
            # every tile published its inclusive word (flag 2) in its own 128-byte slot
            status = words[64:].reshape(-1, 32)
            assert status.shape[0] == nbytes // ceiling.granule(kind)
            assert np.all(status[:, 1] == 2), "a tile's status word is missing"


@pytest.mark.parametrize("kind", [ceiling.WRITE_256X1, ceiling.MEMSET], ids=["write_stream", "memset"])
def test_write_probes_cover_their_range(kind):
    nbytes = 37 * ceiling.granule(ceiling.WRITE_256X1) * 4
    n = nbytes // 4
    dst = sentinel_u32(n + 256)
    ceiling.run(kind, None, dst, nbytes)
    sync()
    out = from_dev_u32(dst)
    assert np.all(out[n:] == SENTINEL), "wrote past the range"
    if kind == ceiling.MEMSET:
        assert np.all(out[:n] == 0x5BD1E995)
    else:
        # each uint4 holds (its vector index, high word, vector slot, constant)
        v = out[:n].reshape(-1, 4)
        assert np.array_equal(v[:, 0], np.arange(v.shape[0], dtype=np.uint32))
        assert np.all(v[:, 1] == 0) and np.all(v[:, 3] == 0x5BD1E995)


def test_read_probe_leaves_its_sink_alone():
    nbytes = 64 * ceiling.granule(ceiling.READ_1024X1)
    src = to_dev(random_words(nbytes // 4, seed=5))
    sink = sentinel_u32(64)
    ceiling.run(ceiling.READ_1024X1, src, sink, nbytes)
    sync()
    # the sink is written only when a tile's XOR equals one constant
    assert np.all(from_dev_u32(sink)[1:] == SENTINEL)
