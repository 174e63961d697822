"""Image ingest / egress on the GPU, through the C ABI: host interleaved <->
device planes and the fused 8-bit quantisation.  Bars: bit-exact (pure data
movement; quantisation against oracle/color.py's single-rounding form)."""
import numpy as np
import pytest

import oracle.color as OC
import split_amd as sa
from tests.conftest import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("c", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [1, 7, 4096, 513 * 511])
def test_upload_download_round_trip(torch_cuda, c, n):
    torch = torch_cuda
    say(f"io: upload/download c={c} n={n}")
    rng = np.random.default_rng(n * 10 + c)
    host = rng.standard_normal((n, c)).astype(np.float32)
    planes = torch.empty((c, n), dtype=torch.float32, device="cuda")
    sa.make_device_image(host, planes)
    assert np.array_equal(planes.cpu().numpy(), OC.deinterleave(host, c))
    back = np.empty_like(host)
    sa.make_host_image(planes, back)
    assert np.array_equal(back, host)


def test_upload_from_pinned_tensor(torch_cuda):
    torch = torch_cuda
    say("io: pinned host tensor")
    h, w = 300, 200
    host = torch.rand((h, w, 3), dtype=torch.float32).pin_memory()
    planes = torch.empty((3, h * w), dtype=torch.float32, device="cuda")
    sa.make_device_image(host, planes)
    assert torch.equal(planes.cpu(), host.reshape(-1, 3).t().contiguous())
    out = torch.empty((h, w, 3), dtype=torch.float32).pin_memory()
    sa.make_host_image(planes, out)
    assert torch.equal(out, host)


@pytest.mark.parametrize("c", [1, 3, 4])
@pytest.mark.parametrize("n", [5, 4096, 100_003])
def test_device_interleave_forms(torch_cuda, c, n):
    torch = torch_cuda
    say(f"io: device (de)interleave c={c} n={n}")
    hwc = torch.rand((n, c), device="cuda")
    planes = torch.empty((c, n), device="cuda")
    sa.deinterleave(hwc, planes)
    assert torch.equal(planes, hwc.t().contiguous())
    back = torch.empty_like(hwc)
    sa.interleave(planes, back)
    assert torch.equal(back, hwc)


@pytest.mark.parametrize("c", [1, 3, 4])
@pytest.mark.parametrize("n", [3, 4096, 65_537])
def test_quantize_matches_oracle(torch_cuda, c, n):
    torch = torch_cuda
    say(f"io: quantize c={c} n={n}")
    rng = np.random.default_rng(n + c)
    v = rng.uniform(-0.2, 1.2, (c, n)).astype(np.float32)
    # values at the rounding edges: (m - 0.5) / 255 and its neighbours
    m = rng.integers(1, 256, size=min(n, 64))
    e = (m - 0.5) / 255.0
    pick = np.stack([np.float32(e), np.nextafter(np.float32(e), 2),
                     np.nextafter(np.float32(e), -2)]).ravel()[:n]
    v[0, :pick.size] = pick
    V = torch.from_numpy(v).cuda()
    out = torch.empty((n, c), dtype=torch.uint8, device="cuda")
    sa.quantize(V, out)
    assert np.array_equal(out.cpu().numpy(), OC.quantize(v))


def test_io_errors(torch_cuda):
    torch = torch_cuda
    say("io: error paths")
    L = sa.lib()
    planes = torch.empty((3, 16), device="cuda")
    host = np.zeros((16, 3), np.float32)
    t = torch.empty(8, dtype=torch.uint8, device="cuda")
    assert L.split_hip_upload_interleaved(host.ctypes.data, 16, 3, planes.data_ptr(),
                                          t.data_ptr(), 8, None) == 2
    assert L.split_hip_deinterleave(planes.data_ptr(), 16, 5, planes.data_ptr(), None) == 1
    assert L.split_hip_interleaved_workspace(16, 0) == 0
