"""The C++ drop-in interface run for real: a C++ program built against
include/split/... and libsplit_hip.so (tests/dropin/main_sequence.cpp) runs
the reference sample's sequence (samples/src/main.cu:228-310) —
uniform_random_initialize, cluster_workspace + one temp block, cluster,
connected_components, compress_labels, calculate_centroids of the clusters
and of the segments (K = number of segments), propagate_centroids, and
segment means with K = N (a 1-pixel checkerboard at 4-connectivity) in the
same temp — and the oracle checks every output.

Bars: seeded init = the sampler's pixels; centroids within 1e-5 of the FP64
Lloyd trajectory (512x512) or of the FP64 means of the labels (full size);
labels = the FP64 argmin against the previous iteration's centroids except
near ties (< 1e-5 relative), or a dense assignment against the centroids of
an (I - 1)-iteration run; segments bit-exact against the CPU union-find and
its dense rank; means within 1e-5 of FP64; propagation an exact gather.
The drop-in cluster() takes no image width: it walks the points as rows of
16, and its results must equal (bit for bit) the image-tiled cluster of the
same image — the integer sums and exact argmin make the trajectory
independent of the traversal."""
import subprocess

import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

BIN = sa.LIB_DIR / "split_dropin_sequence"
CENT_TOL = 1e-5
TIE_REL = 1e-5


def run_sequence(tmp_path, img, w, h, k, seed, max_iter, threshold, conn, tag):
    if not BIN.exists():
        raise sa.SplitLibraryError(f"{BIN} is missing: run `python -m split_amd.build`")
    inp = tmp_path / f"{tag}.in.bin"
    np.ascontiguousarray(img, dtype=np.float32).tofile(inp)
    out = tmp_path / tag
    say(f"drop-in sequence {tag}: {w}x{h} k={k} max_iter={max_iter} "
        f"threshold={threshold} conn={conn}")
    p = subprocess.run([str(BIN), str(inp), str(w), str(h), str(k), str(seed), str(max_iter),
                        repr(float(threshold)), str(conn), str(out)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    n = w * h
    it, shift, nseg, nchk, temp_bytes = open(f"{out}.txt").read().split()
    nseg, nchk = int(nseg), int(nchk)

    def rd(name, dtype, shape):
        return np.fromfile(f"{out}.{name}.bin", dtype=dtype).reshape(shape)

    return dict(
        iterations=int(it), shift=float(shift), nseg=nseg, nchk=nchk,
        temp_bytes=int(temp_bytes),
        init=rd("init", np.float32, (3, k)), centroids=rd("centroids", np.float32, (3, k)),
        labels=rd("labels", np.int32, (n,)), segments=rd("segments", np.int32, (n,)),
        cluster_means=rd("cluster_means", np.float32, (3, k)),
        segment_means=rd("segment_means", np.float32, (3, nseg)),
        propagated=rd("propagated", np.float32, (3, n)),
        checker=rd("checker", np.int32, (n,)),
        checker_means=rd("checker_means", np.float32, (3, nchk)))


def check_after_cluster(img, w, h, k, conn, r):
    """Everything the sequence computes after cluster(), from its labels."""
    n = w * h
    lab, seg = r["labels"], r["segments"]
    rep, ncomp = oracle.cc_uf(lab.reshape(h, w), conn)
    want, cnt = oracle.compress(rep)
    assert r["nseg"] == cnt == ncomp
    assert np.array_equal(seg, want)
    m, _ = oracle.centroids_f64(lab, img, k)
    assert np.abs(r["cluster_means"] - m).max() <= CENT_TOL
    m, _ = oracle.centroids_f64(seg, img, r["nseg"])
    assert np.abs(r["segment_means"] - m).max() <= CENT_TOL
    assert np.array_equal(r["propagated"], r["segment_means"][:, seg])
    # K = N: every pixel of the checkerboard is its own 4-connected component
    assert r["nchk"] == n
    assert np.array_equal(r["checker"], np.arange(n, dtype=np.int32))
    assert np.abs(r["checker_means"] - img).max() <= CENT_TOL


def test_dropin_sequence_plain_512(torch_cuda, tmp_path):
    w = W.PLAIN_512
    img = W.make_image(w.image)
    r = run_sequence(tmp_path, img, 512, 512, w.k, w.init_seed, 20, -1.0, 4, "plain_512")
    idx = oracle.sample_indices(w.init_seed, w.n, w.k)
    assert np.array_equal(r["init"], img[:, idx])
    assert r["iterations"] == 20
    c20, _ = oracle.lloyd_f64(img, r["init"], 20)
    assert np.abs(r["centroids"] - c20).max() <= CENT_TOL
    c19, _ = oracle.lloyd_f64(img, r["init"], 19)
    l19, gap = oracle.label_f64(img, c19.astype(np.float32))
    assert not ((r["labels"] != l19) & (gap >= TIE_REL)).any()
    check_after_cluster(img, 512, 512, w.k, 4, r)


@pytest.mark.parametrize("w,h,k", [(333, 517, 12), (97, 61, 5)])
def test_dropin_cluster_equals_image_cluster(torch_cuda, tmp_path, w, h, k):
    """W*H not a multiple of 16: the drop-in's rows of 16 end in a partial
    row.  Its trajectory equals the image-tiled cluster's bit for bit."""
    torch = torch_cuda
    spec = W.ImageSpec(w, h, seed=0xD809 + w, sites=max(4, w * h // 3000), shading=True,
                       noise_sigma=0.01)
    img = W.make_image(spec)
    seed = 0xD0 + k
    r = run_sequence(tmp_path, img, w, h, k, seed, 60, 1e-4, 8, f"ragged_{w}x{h}")
    P = torch.from_numpy(img).cuda()
    C = torch.from_numpy(r["init"]).cuda()
    L = torch.empty(w * h, dtype=torch.int32, device="cuda")
    it, shift = sa.cluster(P, C, L, 60, 1e-4, width=w, height=h)
    assert r["iterations"] == it
    assert np.float32(r["shift"]) == np.float32(shift)
    assert np.array_equal(r["centroids"], C.cpu().numpy())
    assert np.array_equal(r["labels"], L.cpu().numpy())
    check_after_cluster(img, w, h, k, 8, r)


def test_dropin_sequence_shaded_4096(torch_cuda, tmp_path):
    """The 4096x4096 k=32 image: the headline settings (shift < 1e-4 or 100
    iterations, 8-connectivity) equal segment() of the image bit for bit;
    fixed 20 and 21 iterations give labels equal to a dense assignment
    against the 20-iteration centroids and centroids at the FP64 means of
    their labels."""
    torch = torch_cuda
    w = W.SHADED_4096
    img = W.make_image(w.image)
    r = run_sequence(tmp_path, img, 4096, 4096, w.k, w.init_seed, w.max_iter, w.threshold,
                     w.connectivity, "shaded_4096")
    P = torch.from_numpy(img).cuda()
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    assert np.array_equal(C.cpu().numpy(), r["init"])
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    res = sa.segment(P, 4096, 4096, C, L, S, max_iter=w.max_iter, threshold=w.threshold,
                     connectivity=w.connectivity)
    assert res.iterations == r["iterations"]
    assert np.array_equal(C.cpu().numpy(), r["centroids"])
    assert np.array_equal(L.cpu().numpy(), r["labels"])
    assert np.array_equal(S.cpu().numpy(), r["segments"])
    assert res.num_segments == r["nseg"]
    check_after_cluster(img, 4096, 4096, w.k, w.connectivity, r)
    # fixed iteration counts: the dense-assignment and FP64-mean properties
    runs = {it: run_sequence(tmp_path, img, 4096, 4096, w.k, w.init_seed, it, -1.0, 8,
                             f"shaded_4096_{it}") for it in (20, 21)}
    dense = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.label_points(torch.from_numpy(runs[20]["centroids"]).cuda(), P, dense,
                    width=4096, height=4096)
    assert np.array_equal(dense.cpu().numpy(), runs[21]["labels"])
    m, cnt = oracle.centroids_f64(runs[21]["labels"], img, w.k)
    live = cnt > 0
    assert np.abs(runs[21]["centroids"][:, live] - m[:, live]).max() <= CENT_TOL
