"""The fused extrapolation + flux kernel's tile layout (CPU): the library's
tile order (recursive coordinate bisection) and tile metadata, checked by replaying, for every
interior half-edge, the LDS address the kernel forms (anuga_core_amd/csrc/
de_kernels.hip, de_xflux) against the triangle and edge that the address
holds. Calls only host-side entry points of the library."""

import ctypes as C

import numpy as np
import pytest

from anuga_core_amd import _native
from anuga_core_amd.mesh import Mesh, rectangular_cross

MAX_HALO = 320       # above this the library keeps the separate kernels


def _block():
    """Triangles per tile (DE_BLOCK) of the loaded build."""
    b, w = C.c_int64(), C.c_int64()
    assert _native.lib().anuga_build_config(C.byref(b), C.byref(w)) == 0
    return b.value


BLOCK = _block()


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def tile_layout(mesh):
    """(internal packed neighbours, start, halo, slot, hcap) as the library
    builds them for this mesh."""
    lib = _native.lib()
    N = mesh.number_of_triangles
    cen = np.ascontiguousarray(mesh.centroid_coordinates, np.float64)
    order = np.empty(N, np.int64)
    nbr = np.ascontiguousarray(mesh.neighbours, np.int64)
    assert lib.anuga_tile_order(N, cen.ctypes.data_as(C.POINTER(C.c_double)), _i64(nbr),
                                _i64(order)) == 0
    assert np.array_equal(np.sort(order), np.arange(N))
    iperm = np.empty(N, np.int64)
    iperm[order] = np.arange(N)
    nbr = mesh.neighbours[order]
    nbe = mesh.neighbour_edges[order]
    nb = np.where(nbr >= 0, (iperm[np.maximum(nbr, 0)] << 2) | nbe, -1).astype(np.int32)
    nb = np.ascontiguousarray(nb)
    nt = (N + BLOCK - 1) // BLOCK
    start = np.empty(nt + 1, np.int64)
    slot = np.empty(N, np.uint32)
    nhalo, hcap = C.c_int64(), C.c_int64()
    cap = 3 * N
    halo = np.empty(cap, np.int32)
    rc = lib.anuga_tile_halos(N, nb.ctypes.data_as(C.POINTER(C.c_int32)), _i64(start), cap,
                              halo.ctypes.data_as(C.POINTER(C.c_int32)),
                              slot.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(nhalo),
                              C.byref(hcap))
    assert rc == 0
    return nb, start, halo[:nhalo.value], slot, hcap.value


def replay_lds_addresses(nb, start, halo, slot, hcap):
    """For every interior half-edge (k, i): the kernel's LDS address of the
    neighbour's stage value, decoded back to (triangle, edge) through the
    layout the kernel writes; asserts it is in bounds and is (n, e)."""
    N = len(nb)
    stride_h = max((hcap + 31) // 32 * 32, 32)       # the library's rounding
    lds_len = 12 * BLOCK + 12 * stride_h
    k = np.repeat(np.arange(N), 3)
    i = np.tile(np.arange(3), N)
    v = nb.ravel().astype(np.int64)
    inter = v >= 0
    k, i, v = k[inter], i[inter], v[inter]
    n, e = v >> 2, v & 3
    b = k // BLOCK
    k0 = b * BLOCK
    s = (slot[k].astype(np.int64) >> (10 * i)) & 1023
    in_tile = s == 1023
    # kernel: src = lds + e*BLOCK + (n - k0) (tile) or hl + e*hcap + s (halo);
    # quantity q at src + q*stride with stride 3*BLOCK or 3*hcap
    addr = np.where(in_tile, e * BLOCK + (n - k0), 12 * BLOCK + e * stride_h + s)
    stride = np.where(in_tile, 3 * BLOCK, 3 * stride_h)
    for q in range(4):
        a = addr + q * stride
        assert a.min() >= 0 and a.max() < lds_len
    # decode what the kernel wrote at addr (quantity 0)
    tile_part = addr < 12 * BLOCK
    assert np.array_equal(tile_part, in_tile)
    t_edge = addr[tile_part] // BLOCK
    t_tri = k0[tile_part] + addr[tile_part] % BLOCK
    assert np.all(t_tri < N) and np.all(t_tri // BLOCK == b[tile_part])
    assert np.array_equal(t_tri, n[tile_part]) and np.array_equal(t_edge, e[tile_part])
    ha = addr[~tile_part] - 12 * BLOCK
    h_edge = ha // stride_h
    j = ha % stride_h
    nh = start[b[~tile_part] + 1] - start[b[~tile_part]]
    assert np.all(j < nh)
    entry = halo[start[b[~tile_part]] + j].astype(np.int64)
    assert np.array_equal(entry, n[~tile_part]) and np.array_equal(h_edge, e[~tile_part])
    # the halo holds exactly the neighbour triangles outside each tile, once
    nt = len(start) - 1
    for t in range(nt):
        h = halo[start[t]:start[t + 1]]
        sel = (~in_tile) & (b == t)
        assert len(np.unique(h)) == len(h)
        assert np.array_equal(np.sort(h), np.unique(n[sel]))
    return len(halo)


@pytest.mark.parametrize("m,n", [(200, 200), (37, 53), (1, 1), (8, 8)])
def test_lds_addresses_of_every_half_edge(m, n):
    mesh = Mesh(*rectangular_cross(m, n, 1000.0, 700.0))
    nb, start, halo, slot, hcap = tile_layout(mesh)
    nhalo = replay_lds_addresses(nb, start, halo, slot, hcap)
    assert hcap <= MAX_HALO
    nt = len(start) - 1
    if m == 200:
        # compact tiles: about 52 halo triangles per 256 triangles on this mesh
        assert nhalo / nt < 60 * (BLOCK / 256) ** 0.5 + 8 and hcap < 160
    # neighbours are mutual in the internal numbering
    v = nb.ravel().astype(np.int64)
    kk = np.repeat(np.arange(len(nb)), 3)[v >= 0]
    v = v[v >= 0]
    assert np.array_equal(nb[v >> 2, v & 3] >> 2, kk)


def test_tile_order_is_local():
    """Consecutive triangles in the tile order are close, and each run of 256
    is a compact patch of about 8 x 8 cells."""
    mesh = Mesh(*rectangular_cross(100, 100, 100.0, 100.0))
    nb, start, halo, slot, hcap = tile_layout(mesh)
    lib = _native.lib()
    cen = np.ascontiguousarray(mesh.centroid_coordinates)
    order = np.empty(len(cen), np.int64)
    nbr = np.ascontiguousarray(mesh.neighbours, np.int64)
    lib.anuga_tile_order(len(cen), cen.ctypes.data_as(C.POINTER(C.c_double)), _i64(nbr),
                         _i64(order))
    c = cen[order]
    step = np.linalg.norm(c[1:] - c[:-1], axis=1)
    assert np.median(step) < 1.0            # cells are 1 m
    spread = [np.ptp(c[t:t + BLOCK], axis=0).max() for t in range(0, len(c) - BLOCK, BLOCK)]
    assert np.median(spread) <= 16.0 * (BLOCK / 256) ** 0.5   # 256 triangles: about 8 cells


def test_tile_order_is_deterministic():
    """The bisection cuts parts on several threads; the order is the same on
    every call, and a permutation, with and without the neighbours."""
    mesh = Mesh(*rectangular_cross(300, 240, 300.0, 240.0))
    lib = _native.lib()
    cen = np.ascontiguousarray(mesh.centroid_coordinates)
    nbr = np.ascontiguousarray(mesh.neighbours, np.int64)
    N = len(cen)
    runs = []
    for with_nbr in (True, True, False):
        order = np.empty(N, np.int64)
        assert lib.anuga_tile_order(N, cen.ctypes.data_as(C.POINTER(C.c_double)),
                                    _i64(nbr) if with_nbr else None, _i64(order)) == 0
        assert np.array_equal(np.sort(order), np.arange(N))
        runs.append(order)
    assert np.array_equal(runs[0], runs[1])
    bad = nbr.copy()
    bad[0, 0] = N
    order = np.empty(N, np.int64)
    assert lib.anuga_tile_order(N, cen.ctypes.data_as(C.POINTER(C.c_double)), _i64(bad),
                                _i64(order)) != 0


def test_tile_halos_are_short():
    """The bisection's runs of 256 triangles are compact: on a 200 x 200
    cross mesh the mean halo (neighbours outside the tile) is under 40
    triangles (an 8 x 8-cell patch cut along cell lines has 32), and no tile
    needs a second halo wave (at most 64). The halo is the work the fused
    kernel repeats per tile."""
    if BLOCK != 256:
        pytest.skip("sized for 256-triangle tiles")
    mesh = Mesh(*rectangular_cross(200, 200, 200.0, 200.0))
    nb, start, halo, slot, hcap = tile_layout(mesh)
    nh = np.diff(start)
    assert nh.mean() < 40.0, nh.mean()
    assert nh.max() <= 64, nh.max()
    # every tile's halo is exactly the outside neighbours of its triangles
    for t in range(0, len(nh), 37):
        k0, k1 = t * BLOCK, min((t + 1) * BLOCK, len(nb))
        v = nb[k0:k1].ravel()
        n = v[v >= 0] >> 2
        outside = np.unique(n[(n < k0) | (n >= k1)])
        assert np.array_equal(np.sort(halo[start[t]:start[t + 1]]), outside)


def test_tile_halos_rejects_bad_neighbours():
    lib = _native.lib()
    nb = np.array([[4 * 5, -1, -1]], np.int32)           # neighbour 5 of a 1-triangle mesh
    start = np.empty(2, np.int64)
    halo = np.empty(3, np.int32)
    slot = np.empty(1, np.uint32)
    a, b = C.c_int64(), C.c_int64()
    rc = lib.anuga_tile_halos(1, nb.ctypes.data_as(C.POINTER(C.c_int32)), _i64(start), 3,
                              halo.ctypes.data_as(C.POINTER(C.c_int32)),
                              slot.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(a), C.byref(b))
    assert rc != 0


def tile_edge_lists(nb):
    lib = _native.lib()
    N = len(nb)
    nt = (N + BLOCK - 1) // BLOCK
    estart = np.empty(nt + 1, np.int64)
    esplit = np.empty(nt, np.int64)
    cap = 3 * N + 1
    edges = np.empty(cap, np.uint32)
    ne = C.c_int64()
    rc = lib.anuga_tile_edges(N, nb.ctypes.data_as(C.POINTER(C.c_int32)), _i64(estart),
                              _i64(esplit), cap, edges.ctypes.data_as(C.POINTER(C.c_uint32)),
                              C.byref(ne))
    assert rc == 0
    edges = edges[:ne.value]
    # each tile's edges inside the tile come first, then its edges to the halo
    for b in range(nt):
        assert estart[b] <= esplit[b] <= estart[b + 1]
        halo = (edges[estart[b]:estart[b + 1]] >> 12) & 1
        assert np.all(halo[:esplit[b] - estart[b]] == 0) and np.all(halo[esplit[b] - estart[b]:] == 1)
    return estart, edges


def replay_edge_once(nb, start, halo, slot, hcap, estart, edges):
    """The once-per-edge kernel's LDS dataflow (de_xflux_e) replayed with
    tokens: phase 1 writes each edge value as (k, edge, quantity); phase 2a
    must read, for each listed edge, its own and its neighbour's values, and
    its writes must not touch a slot another entry reads or writes; phase 2b
    must read, for every half-edge, the terms its own edge produced, through
    the roles (this side, other side, halo, boundary) the kernel decodes.
    Returns the number of entries."""
    N = len(nb)
    hs = max((hcap + 31) // 32 * 32, 32)
    Q = 3 * BLOCK
    size = 14 * BLOCK + 12 * hs
    nt = len(start) - 1
    listed = set()
    for b in range(nt):
        k0, k1 = b * BLOCK, min(N, (b + 1) * BLOCK)
        lds = {}
        for k in range(k0, k1):
            for i in range(3):
                for q in range(4):
                    lds[(3 * q + i) * BLOCK + (k - k0)] = ("val", k, i, q)
        hb = halo[start[b]:start[b + 1]]
        for s, h in enumerate(hb):
            for i in range(3):
                for q in range(4):
                    lds[12 * BLOCK + (3 * q + i) * hs + s] = ("val", int(h), i, q)
        touched = set()
        # the kernel's split of the list: in phase 2 each thread of the waves
        # not on the halo takes one entry inside the tile, phase 3 the rest
        e0, e1 = int(estart[b]), int(estart[b + 1])
        es = e0 + int(np.sum(((edges[e0:e1] >> 12) & 1) == 0))
        hw = min((len(hb) + 63) // 64, BLOCK // 64 - 1)
        p2 = list(range(e0, min(es, e0 + BLOCK - 64 * hw)))
        p3 = list(range(min(es, e0 + BLOCK - 64 * hw), e1))
        assert p2 + p3 == list(range(e0, e1))
        assert all(((int(edges[e]) >> 12) & 1) == 0 for e in p2)
        for en in edges[e0:e1]:
            en = int(en)
            t, i, isha, nn, j = en & 1023, (en >> 10) & 3, (en >> 12) & 1, (en >> 13) & 1023, \
                (en >> 23) & 3
            k = k0 + t
            assert k < k1
            v = int(nb[k, i])
            assert v >= 0 and (v & 3) == j
            n = v >> 2
            A = i * BLOCK + t
            if isha:
                assert nn < len(hb) and int(hb[nn]) == n
                R, rs = 12 * BLOCK + j * hs + nn, 3 * hs
            else:
                assert k0 <= n < k1 and n - k0 == nn and k < n
                R, rs = j * BLOCK + nn, Q
            addrs = [A + q * Q for q in range(4)] + [R + q * rs for q in range(4)]
            for a in addrs:
                assert 0 <= a < 12 * BLOCK + 12 * hs
                assert a not in touched, "two entries share an LDS slot"
                # an edge inside the tile touches only the tile's values: it runs
                # while the halo's are being written
                assert isha or a < 12 * BLOCK
            for q in range(4):
                assert lds[A + q * Q] == ("val", k, i, q)
                assert lds[R + q * rs] == ("val", n, j, q)
            touched.update(addrs)
            assert (k, i) not in listed
            listed.add((k, i))
            for q, name in enumerate(("a0", "a1", "a2", "b1")):
                lds[A + q * Q] = (name, k, i)
            lds[R] = ("b2", k, i)
            if not isha:
                lds[R + rs] = ("c1", k, i)
                lds[R + 2 * rs] = ("c2", k, i)
        # phase 2b
        for k in range(k0, k1):
            t = k - k0
            for i in range(3):
                v = int(nb[k, i])
                A = i * BLOCK + t
                if v < 0:
                    assert all(lds[A + q * Q] == ("val", k, i, q) for q in range(4))
                    continue
                n, j = v >> 2, v & 3
                s = (int(slot[k]) >> (10 * i)) & 1023
                if s != 1023:
                    got = [lds[A + q * Q] for q in range(4)] + [lds[12 * BLOCK + j * hs + s]]
                    want = [(x, k, i) for x in ("a0", "a1", "a2", "b1", "b2")]
                elif k < n:
                    got = [lds[A + q * Q] for q in range(4)] + [lds[j * BLOCK + (n - k0)]]
                    want = [(x, k, i) for x in ("a0", "a1", "a2", "b1", "b2")]
                else:
                    O = j * BLOCK + (n - k0)
                    got = [lds[O + q * Q] for q in range(3)] + [lds[A + Q], lds[A + 2 * Q]]
                    want = [(x, n, j) for x in ("a0", "a1", "a2", "c1", "c2")]
                assert got == want, (k, i, got, want)
    # every half-edge with an interior neighbour is listed once from one side
    for k in range(N):
        for i in range(3):
            v = int(nb[k, i])
            if v < 0:
                continue
            n = v >> 2
            same_tile = n // BLOCK == k // BLOCK
            assert ((k, i) in listed) == (not same_tile or k < n)
    return len(edges)


@pytest.mark.parametrize("m,n", [(40, 40), (37, 53), (1, 1), (8, 8)])
def test_edge_once_lds_dataflow(m, n):
    mesh = Mesh(*rectangular_cross(m, n, 1000.0, 700.0))
    nb, start, halo, slot, hcap = tile_layout(mesh)
    estart, edges = tile_edge_lists(nb)
    ne = replay_edge_once(nb, start, halo, slot, hcap, estart, edges)
    N = len(nb)
    if m == 40:
        # about 1.6 evaluated edges per triangle instead of 3
        assert 1.4 < ne / N < 1.8


def test_edge_once_dataflow_on_unstructured_mesh():
    from anuga_core_amd.mesh import graded_square
    mesh = Mesh(*graded_square(600.0, 40.0, 300.0, (150.0, 150.0, 450.0, 450.0), seed=1))
    nb, start, halo, slot, hcap = tile_layout(mesh)
    estart, edges = tile_edge_lists(nb)
    replay_edge_once(nb, start, halo, slot, hcap, estart, edges)
