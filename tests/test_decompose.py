"""Droplet labelling (gp::decompose::decompose, src/cluster.cpp:17-82; exposed as
gpCpp.decomposition, src/gpCpp.cpp:289-313).

The oracle's restatement is pinned on the CPU against hand-built cases
(26-neighbour connectivity, periodic wrap, scan-order numbering, the strict
cut-off), against scipy.ndimage.label away from the boundary, and against a
brute-force periodic union-find on tiny grids.  The GPU labels must equal the
oracle's exactly."""
import itertools

import numpy as np
import pytest

from conftest import seeded_field


def _field_from_density(rho):
    rho = np.asarray(rho, dtype=float)
    if rho.ndim == 3:
        rho = rho[..., None]
    return np.asfortranarray(np.sqrt(rho).astype(complex))


def _scan_order_relabel(lab):
    """Renumber positive labels by first occurrence with k fastest, i slowest."""
    out = np.where(lab > 0, 0, -1).astype(np.int32)
    nx, ny, nz = lab.shape
    seen = {}
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                v = lab[i, j, k]
                if v > 0:
                    if v not in seen:
                        seen[v] = len(seen) + 1
                    out[i, j, k] = seen[v]
    return out, len(seen)


def _brute_periodic(mask):
    """Union-find over the 26 periodic neighbours, pure Python."""
    nx, ny, nz = mask.shape
    parent = {}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    cells = [c for c in itertools.product(range(nx), range(ny), range(nz)) if mask[c]]
    for c in cells:
        parent[c] = c
    for (i, j, k) in cells:
        for a, b, d in itertools.product((-1, 0, 1), repeat=3):
            n = ((i + a) % nx, (j + b) % ny, (k + d) % nz)
            if mask[n]:
                ra, rb = find((i, j, k)), find(n)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    lab = np.zeros(mask.shape, dtype=np.int64)
    roots = {}
    for c in cells:
        lab[c] = roots.setdefault(find(c), len(roots) + 1)
    return _scan_order_relabel(lab)


# ---------------------------------------------------------------- oracle pins (CPU)

def test_oracle_hand_built_cases(oracle_mod):
    O = oracle_mod
    rho = np.zeros((8, 8, 8))
    rho[5, 1, 1] = rho[5, 1, 2] = 2.0          # droplet with the larger i
    rho[2, 6, 6] = 2.0                          # smaller i: numbered first despite larger j, k
    rho[3, 3, 3] = rho[4, 4, 4] = 2.0          # corner contact: one droplet
    rho[0, 4, 0] = rho[7, 4, 0] = 2.0          # periodic in x
    rho[6, 0, 7] = 1.0                          # exactly at the cut-off: background
    labels, n = O.decompose(_field_from_density(rho), 1.0)
    L = labels[..., 0]
    assert n == 4
    assert L[0, 4, 0] == 1 and L[7, 4, 0] == 1
    assert L[2, 6, 6] == 2
    assert L[3, 3, 3] == L[4, 4, 4] == 3
    assert L[5, 1, 1] == L[5, 1, 2] == 4
    assert L[6, 0, 7] == -1
    assert (L == -1).sum() == 512 - 7


def test_oracle_periodic_corner_and_all_or_nothing(oracle_mod):
    O = oracle_mod
    rho = np.zeros((4, 5, 6))
    rho[0, 0, 0] = rho[3, 4, 5] = 3.0           # opposite corners touch through the wrap
    labels, n = O.decompose(_field_from_density(rho), 0.5)
    assert n == 1 and labels[0, 0, 0, 0] == labels[3, 4, 5, 0] == 1
    labels, n = O.decompose(_field_from_density(np.ones((3, 3, 3))), 0.5)
    assert n == 1 and (labels == 1).all()
    labels, n = O.decompose(_field_from_density(np.ones((3, 3, 3))), 1.0)
    assert n == 0 and (labels == -1).all()


def test_oracle_matches_scipy_label_inside_the_box(oracle_mod):
    ndimage = pytest.importorskip("scipy.ndimage")
    O = oracle_mod
    rng = np.random.default_rng(4)
    rho = rng.uniform(0, 1, (20, 18, 16))
    rho[0, :, :] = rho[-1, :, :] = 0
    rho[:, 0, :] = rho[:, -1, :] = 0
    rho[:, :, 0] = rho[:, :, -1] = 0
    cut = 0.8
    lab, _ = ndimage.label(rho > cut, structure=np.ones((3, 3, 3)))
    want, nwant = _scan_order_relabel(lab)
    got, n = O.decompose(_field_from_density(rho), cut)
    assert n == nwant > 20
    assert np.array_equal(got[..., 0], want)


@pytest.mark.parametrize("shape,fill,seed", [((5, 4, 3), 0.3, 0), ((6, 6, 6), 0.15, 1), ((1, 7, 9), 0.4, 2),
                                             ((2, 2, 8), 0.3, 3), ((9, 1, 1), 0.5, 4)])
def test_oracle_matches_brute_force_periodic(oracle_mod, shape, fill, seed):
    O = oracle_mod
    rho = np.random.default_rng(seed).uniform(0, 1, shape)
    mask = rho > 1 - fill
    want, nwant = _brute_periodic(mask)
    got, n = O.decompose(_field_from_density(rho), 1 - fill)
    assert n == nwant
    assert np.array_equal(got[..., 0], want)


def test_oracle_labels_only_the_requested_component(oracle_mod):
    O = oracle_mod
    psi = seeded_field((6, 5, 4), 2, seed=9)
    labels, n = O.decompose(psi, 0.5, component=1)
    assert (labels[..., 0] == 0).all()
    assert ((labels[..., 1] == -1) | (labels[..., 1] >= 1)).all()
    assert n == labels[..., 1].max()


# ---------------------------------------------------------------- GPU ---------

def _blobs(shape, count, seed, radius=3.0):
    """Density with Gaussian droplets at random centres (periodic images included)."""
    rng = np.random.default_rng(seed)
    axes = [np.arange(n, dtype=float) for n in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    rho = np.zeros(shape)
    for _ in range(count):
        c = rng.uniform(0, 1, 3) * np.array(shape)
        d2 = 0
        for A, cc, n in zip((X, Y, Z), c, shape):
            d = np.abs(A - cc)
            d = np.minimum(d, n - d)
            d2 = d2 + d * d
        rho += rng.uniform(0.5, 2.0) * np.exp(-d2 / (2 * radius ** 2))
    return rho


@pytest.mark.gpu
@pytest.mark.parametrize("shape,kind,seed", [
    ((16, 16, 16), "noise", 0), ((17, 9, 5), "noise", 1), ((1, 7, 9), "noise", 2), ((2, 2, 8), "noise", 3),
    ((64, 64, 64), "blobs", 4), ((96, 80, 72), "blobs", 5), ((128, 128, 128), "noise", 6),
    ((30, 30, 30), "all", 7), ((30, 30, 30), "none", 8),
])
def test_gpu_labels_equal_oracle(oracle_mod, shape, kind, seed):
    from gp5_amd import capi

    O = oracle_mod
    if kind == "noise":
        rho = np.random.default_rng(seed).uniform(0, 1, shape)
        cut = 0.75
    elif kind == "blobs":
        rho = _blobs(shape, 24, seed)
        cut = 0.3
    elif kind == "all":
        rho, cut = np.ones(shape), 0.5
    else:
        rho, cut = np.ones(shape), 1.5
    psi = _field_from_density(rho) * np.exp(0.3j)
    print(f"decompose shape={shape} kind={kind} seed={seed}", flush=True)
    ctx = capi.Context(shape, [-1] * 3, [1] * 3, 1)
    buf = ctx.field(psi)
    got, n = capi.decompose(ctx, buf, cut)
    want, nwant = O.decompose(psi, cut)
    assert n == nwant
    assert np.array_equal(got, want)
    # a second call reuses the workspace
    got2, n2 = capi.decompose(ctx, buf, cut)
    assert n2 == n and np.array_equal(got2, got)


@pytest.mark.gpu
def test_gpu_labels_second_component_and_pybind(oracle_mod):
    from gp5_amd import capi, gp, gpCpp

    O = oracle_mod
    shape = (24, 20, 16)
    rho0 = _blobs(shape, 6, 11)
    rho1 = _blobs(shape, 9, 12, radius=2.0)
    psi = np.asfortranarray(np.stack([np.sqrt(rho0), np.sqrt(rho1)], axis=3).astype(complex))
    ctx = capi.Context(shape, [-1] * 3, [1] * 3, 2)
    buf = ctx.field(psi)
    got, n = capi.decompose(ctx, buf, 0.25, comp=1)
    want, nwant = O.decompose(psi, 0.25, component=1)
    assert n == nwant and np.array_equal(got, want)
    # gpCpp.decomposition labels component 0, as the reference's binding does
    geo = gp.geometry(shape, left=[-1] * 3, right=[1] * 3)
    f = gp.field(geo, nComponents=2)
    f.array = psi
    lab = gpCpp.decomposition().decompose(f._fieldCpp, 0.25)
    want0, _ = O.decompose(psi, 0.25, component=0)
    assert lab.shape == shape + (2,) and lab.dtype == np.int32
    assert np.array_equal(lab, want0)
