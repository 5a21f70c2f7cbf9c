"""torch.ops.gps.* registration, fake (meta) shapes and argument errors; no GPU."""
import pytest

torch = pytest.importorskip("torch")

import gpuprefixsums_amd.torch_ops  # noqa: E402,F401  (registers the operators)


@pytest.mark.parametrize("dtype,out", [(torch.uint32, torch.uint32), (torch.int32, torch.int32),
                                       (torch.float32, torch.float32), (torch.bfloat16, torch.float32)])
def test_fake_shapes_and_dtypes(dtype, out):
    x = torch.empty(1000, dtype=dtype, device="meta")
    y = torch.ops.gps.cumsum(x)
    assert y.shape == x.shape and y.dtype == out and y.device.type == "meta"
    off = torch.empty(5, dtype=torch.int64, device="meta")
    s = torch.ops.gps.segmented_cumsum(x, off, inclusive=dtype not in (torch.uint32, torch.int32))
    assert s.shape == x.shape and s.dtype == out


def test_fake_tensor_mode_traces_without_a_device():
    from torch._subclasses.fake_tensor import FakeTensorMode
    with FakeTensorMode():
        z = torch.empty(33, dtype=torch.bfloat16, device="cuda")
        assert torch.ops.gps.cumsum(z).dtype == torch.float32


def test_cpu_tensors_and_bad_arguments_are_rejected():
    with pytest.raises(RuntimeError, match="ROCm device"):
        torch.ops.gps.cumsum(torch.zeros(4, dtype=torch.float32))
    with pytest.raises(RuntimeError, match="ROCm device"):
        torch.ops.gps.segmented_cumsum(torch.zeros(4, dtype=torch.int32), torch.zeros(2, dtype=torch.int64))
