"""Host layout (csrc/layout.cpp): the batched patch layout covers exactly the
brute-force pair set of every system (no missing, extra, duplicate or
cross-system pairs), with the pair kernel's FP32 shift arithmetic, so image
shifts are exact for box edges that are not multiples of the patch grid's
quantum.  The helper is compiled from the library's own layout.cpp."""
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial import cKDTree

from mi355nb import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def helper(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    exe = str(d / "layout_pairs")
    csrc = os.path.join(ROOT, "namd-3.0_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", csrc, os.path.join(ROOT, "tests", "layout_pairs.cpp"),
                    os.path.join(csrc, "layout.cpp"), "-o", exe], check=True)
    return d, exe


def test_batched_layout_covers_exact_pair_sets(helper):
    d, exe = helper
    members = systems.batch_systems(n_systems=3, seed=11, n_min=3000, n_max=5000)
    sysl = [m.system for m in members]
    starts = np.zeros(len(sysl) + 1, np.int32)
    for k, s in enumerate(sysl):
        starts[k + 1] = starts[k] + s.natoms
    ptr = [np.zeros(1, np.int64)]
    idx = []
    off = 0
    for k, s in enumerate(sysl):
        ptr.append(np.asarray(s.excl_ptr[1:], np.int64) + off)
        idx.append(np.asarray(s.excl_idx, np.int32) + starts[k])
        off += int(s.excl_ptr[-1])
    with open(d / "in.bin", "wb") as f:
        np.array([len(sysl)], np.int32).tofile(f)
        starts.tofile(f)
        np.array([s.box for s in sysl], np.float64).tofile(f)
        np.concatenate([s.pos for s in sysl]).astype(np.float64).tofile(f)
        np.concatenate(ptr).tofile(f)
        np.concatenate(idx).astype(np.int32).tofile(f)
    subprocess.run([exe, str(d / "in.bin"), "12.0", "13.5", str(d / "pairs.bin"), str(d / "eff.bin")], check=True,
                   capture_output=True)
    got = np.fromfile(d / "pairs.bin", np.int32).reshape(-1, 2)
    eff = np.fromfile(d / "eff.bin", np.float64).reshape(-1, 3)
    gs = set(map(tuple, got.tolist()))
    assert len(gs) == len(got), "a pair is covered twice"
    for k, s in enumerate(sysl):
        a, b = int(starts[k]), int(starts[k + 1])
        assert (s.box != np.round(s.box * 1024) / 1024).any(), "box edges should not be multiples of 2^-10 A here"
        bq = np.round(s.box * 131072.0) / 131072.0
        p = eff[a:b]
        pr = cKDTree(np.mod(p, bq), boxsize=bq).query_pairs(12.5, output_type="ndarray")
        dd = p[pr[:, 0]] - p[pr[:, 1]]
        dd -= bq * np.round(dd / bq)
        pr = pr[(dd * dd).sum(1) <= 144.0]
        excl = set()
        for i in range(s.natoms):
            for j in s.excl_idx[s.excl_ptr[i]:s.excl_ptr[i + 1]]:
                excl.add((i, int(j)))
        ref = set((int(i) + a, int(j) + a) for i, j in pr.tolist() if (i, j) not in excl)
        mine = set(x for x in gs if a <= x[0] < b)
        assert all(a <= x[1] < b for x in mine), "cross-system pair"
        assert mine == ref, (k, len(mine - ref), len(ref - mine))


@pytest.mark.parametrize("cell", [
    [[34.0, 0.0, 0.0], [8.0, 33.0, 0.0], [-6.0, 7.0, 34.0]],          # skewed, 2 patches per direction
    [[29.3, 0.0, 0.0], [14.1, 27.7, 0.0], [13.9, 9.2, 28.1]],         # strongly skewed, 1-2 patches
    [[45.1, 1.3, -0.7], [-3.2, 41.9, 2.2], [5.5, -4.1, 47.3]],        # general orientation, 3 patches
])
def test_general_cell_layout_covers_exact_pair_sets(helper, cell):
    """One system in a general cell: the skeleton covers exactly the pairs
    within the cutoff of a brute-force enumeration over the 27 images around
    the fractional minimum image (same rounded lattice, effective positions)."""
    d, exe = helper
    cell = np.asarray(cell)
    base = systems.small_box(n_waters=500, n_chains=2, chain_len=20, box=30.0, seed=8)
    s = systems.sheared(base, cell)
    with open(d / "cell.bin", "wb") as f:
        np.array([-1], np.int32).tofile(f)
        np.array([0, s.natoms], np.int32).tofile(f)
        cell.astype(np.float64).tofile(f)
        s.pos.astype(np.float64).tofile(f)
        np.asarray(s.excl_ptr, np.int64).tofile(f)
        np.asarray(s.excl_idx, np.int32).tofile(f)
    out = subprocess.run([exe, str(d / "cell.bin"), "12.0", "13.5", str(d / "cpairs.bin"), str(d / "ceff.bin")],
                         check=True, capture_output=True, text=True)
    got = np.fromfile(d / "cpairs.bin", np.int32).reshape(-1, 2)
    eff = np.fromfile(d / "ceff.bin", np.float64).reshape(-1, 3)
    gs = set(map(tuple, got.tolist()))
    assert len(gs) == len(got), "a pair is covered twice"
    lat = np.round(cell * 131072.0) / 131072.0
    inv = np.linalg.inv(lat)
    imgs = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], np.float64) @ lat
    excl = set()
    for i in range(s.natoms):
        for j in s.excl_idx[s.excl_ptr[i]:s.excl_ptr[i + 1]]:
            excl.add((min(i, int(j)), max(i, int(j))))
    ref = set()
    for i in range(s.natoms - 1):
        dd = eff[i + 1:] - eff[i]
        dd = dd - np.rint(dd @ inv) @ lat
        r2 = (((dd[:, None, :] + imgs[None, :, :]) ** 2).sum(2) <= 144.0).any(1)
        for j in np.nonzero(r2)[0] + i + 1:
            if (i, int(j)) not in excl:
                ref.add((i, int(j)))
    print(out.stdout.strip())
    assert gs == ref, (len(gs - ref), len(ref - gs))
