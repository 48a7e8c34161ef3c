"""mwd_copy_arrays: several device arrays copied in one launch."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_copy_arrays_copies_every_array(gpu):
    from mwdiem_amd import _lib
    g = torch.Generator(device="cpu").manual_seed(7)
    sizes = [1, 3, 0, 4097, 12, 65536 * 3 + 5]
    src = [torch.randn(n, generator=g, dtype=torch.float64).cuda() for n in sizes]
    src.append(torch.arange(11, dtype=torch.int64).cuda())
    src.append(torch.arange(6, dtype=torch.int32).cuda())  # 24 bytes
    dst = [torch.full_like(s, -7) for s in src]
    _lib.copy_arrays(dst, src)
    torch.cuda.synchronize()
    for d, s in zip(dst, src):
        assert torch.equal(d, s)


def test_copy_arrays_refuses_bad_arguments(gpu):
    from mwdiem_amd import _lib
    a = torch.zeros(4, dtype=torch.float64, device="cuda")
    b = torch.zeros(5, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _lib.copy_arrays([a], [b])
    with pytest.raises(ValueError):
        _lib.copy_arrays([a] * (_lib.MAX_COPY + 1), [a] * (_lib.MAX_COPY + 1))
    odd = torch.zeros(3, dtype=torch.int8, device="cuda")
    with pytest.raises(_lib.MwdError):
        _lib.copy_arrays([odd], [torch.zeros(3, dtype=torch.int8, device="cuda")])
