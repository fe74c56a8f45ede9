"""GPU: the streaming copy the benchmark uses as its bandwidth yardstick
(r2d3c_copy_probe) copies exactly, including a ragged tail, and rejects
sizes that are not a multiple of 16 bytes and unaligned pointers."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, (1 << 22) + 16 * 37])
def test_copy_probe_copies(nbytes, cuda):
    print("copy probe %d bytes" % nbytes, flush=True)
    from rot2d3c_amd import _lib, spectral as S
    g = torch.Generator(device="cpu").manual_seed(nbytes)
    src = torch.randint(0, 255, (nbytes + 64,), dtype=torch.uint8, generator=g).to(cuda)
    dst = torch.zeros_like(src)
    _lib.check(_lib.lib().r2d3c_copy_probe(dst.data_ptr(), src.data_ptr(), nbytes, S._stream(cuda)),
               "copy_probe")
    torch.cuda.synchronize()
    assert torch.equal(dst[:nbytes], src[:nbytes])
    assert int(dst[nbytes:].abs().sum()) == 0  # nothing past the end


def test_copy_probe_rejects_ragged_sizes(cuda):
    from rot2d3c_amd import _lib, spectral as S
    buf = torch.zeros(64, dtype=torch.uint8, device=cuda)
    assert _lib.lib().r2d3c_copy_probe(buf.data_ptr(), buf.data_ptr(), 24, S._stream(cuda)) != 0
    assert _lib.lib().r2d3c_copy_probe(buf.data_ptr(), buf.data_ptr(), -16, S._stream(cuda)) != 0
    assert _lib.lib().r2d3c_copy_probe(buf.data_ptr() + 8, buf.data_ptr(), 16, S._stream(cuda)) != 0
