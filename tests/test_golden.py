"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the generator and the oracle reproduce the frozen vectors bit for bit.
GPU: the gfx950 path reproduces them to the parity bars (labels up to
near ties, centroids within 1e-5 of FP64, components bit-exact)."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from split_amd import workloads as W

FIX = sorted((Path(__file__).parent / "golden").glob("*.npz"))


def load(p):
    return dict(np.load(p, allow_pickle=False))


@pytest.mark.parametrize("path", FIX, ids=lambda p: p.stem)
def test_generator_and_oracle_reproduce_fixture(path):
    f = load(path)
    w, h = int(f["width"]), int(f["height"])
    spec = W.ImageSpec(w, h, seed=int(f["gen_seed"]), sites=int(f["sites"]),
                       shading=bool(f["shading"]), noise_sigma=float(f["sigma"]))
    img = W.make_image(spec)
    assert np.array_equal(img, f["image"])
    idx = oracle.sample_indices(int(f["init_seed"]), w * h, int(f["k"]))
    assert np.array_equal(idx, f["init_index"])
    cent, lab = oracle.lloyd_f64(img, f["init"], int(f["iters"]))
    assert np.array_equal(cent, f["centroids"]) and np.array_equal(lab, f["labels"])
    assert np.array_equal(oracle.ref_label_points(f["init"], img), f["ref_labels_init"])
    assert np.array_equal(oracle.cc_uf(lab.reshape(h, w), 4)[0], f["cc4"])
    assert np.array_equal(oracle.cc_uf(lab.reshape(h, w), 8)[0], f["cc8"])
    assert np.array_equal(oracle.compress(f["cc8"])[0], f["rank8"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIX, ids=lambda p: p.stem)
def test_gpu_matches_fixture(torch_cuda, path):
    import split_amd as sa
    from tests.conftest import say
    torch = torch_cuda
    f = load(path)
    say(f"golden {path.stem}")
    w, h, k, iters = int(f["width"]), int(f["height"]), int(f["k"]), int(f["iters"])
    P = torch.from_numpy(f["image"]).cuda()
    C = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(P, C, int(f["init_seed"]))
    torch.cuda.synchronize()
    assert np.array_equal(C.cpu().numpy(), f["init"])
    L = torch.empty(w * h, dtype=torch.int32, device="cuda")
    sa.label_points(C, P, L)
    torch.cuda.synchronize()
    assert np.array_equal(L.cpu().numpy(), f["ref_labels_init"])
    it, _ = sa.cluster(P, C, L, iters, -1.0)
    assert it == iters
    assert np.abs(C.cpu().numpy() - f["centroids"]).max() <= 1e-5
    # labels of the last assignment against the FP64 centroids one update back
    c_prev, _ = oracle.lloyd_f64(f["image"], f["init"], iters - 1)
    l64, gap = oracle.label_f64(f["image"], c_prev.astype(np.float32))
    got = L.cpu().numpy()
    assert not ((got != f["labels"]) & (gap >= 1e-5)).any()
    # components and ranks of the fixture labels
    lab = torch.from_numpy(f["labels"].reshape(h, w).copy()).cuda()
    for conn, key in ((4, "cc4"), (8, "cc8")):
        out = torch.empty(w * h, dtype=torch.int32, device="cuda")
        sa.connected_components(lab, out, connectivity=conn)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), f[key])
    out = torch.from_numpy(f["cc8"].copy()).cuda()
    assert sa.compress_labels(out) == int(f["ncc8"])
    assert np.array_equal(out.cpu().numpy(), f["rank8"])
