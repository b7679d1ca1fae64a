"""anuga_water_volume_noalloc (the entry point that allocates nothing) against
anuga_get_water_volume on the device (MI355X), through the C ABI: the same
bits for the same state arrays in the same order, on two calls in a row (the
second on a smaller mesh, so the static scratch keeps the first call's
larger partials), and the size limit refused."""

import ctypes as C

import numpy as np
import pytest
import torch

from anuga_core_amd import _native
from anuga_core_amd.domain import Domain
from tests import scenarios as S

pytestmark = pytest.mark.gpu


def _volume_pair(m, n):
    """(anuga_get_water_volume of a mesh-order Domain, the no-allocation
    variant over its stage, elevation and area)."""
    mesh, _, q, bc = S.hills_de1(m=m, n=n)
    d = S.setup(Domain(mesh=mesh, reorder=False), "DE1", q, bc)
    d.evolve_steps(5)
    d.distribute_to_vertices_and_edges()
    ref = d.get_water_volume()
    dev = torch.device("cuda:0")
    arrays = [torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
              for a in (d.get_quantity("stage").centroid_values,
                        d.get_quantity("elevation").centroid_values, mesh.areas)]
    out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    rc = _native.lib().anuga_water_volume_noalloc(
        mesh.number_of_triangles, *[C.c_void_p(a.data_ptr()) for a in arrays],
        C.c_void_p(out.data_ptr()), C.c_void_p(stream.cuda_stream))
    assert rc == 0, _native.lib().anuga_last_error(None)
    stream.synchronize()
    d.close()
    return ref, float(out.item()), mesh.number_of_triangles


def test_noalloc_volume_matches_the_handle_volume_twice():
    print("case no-alloc volume", flush=True)
    big, small = _volume_pair(40, 33), _volume_pair(9, 7)
    assert big[2] > small[2] and big[2] % 256 != 0
    for ref, got, _ in (big, small):
        assert got == ref, (got, ref)
    # the second call again after the first, in the same process
    again = _volume_pair(40, 33)
    assert again[1] == big[1]


def test_noalloc_volume_refuses_too_many_triangles():
    print("case no-alloc refusals", flush=True)
    L = _native.lib()
    p = C.c_void_p(1)
    assert L.anuga_water_volume_noalloc((1 << 27) + 1, p, p, p, p, None) == _native.ANUGA_E_ARG
    assert b"2^27" in L.anuga_last_error(None)
    assert L.anuga_water_volume_noalloc(0, p, p, p, p, None) == _native.ANUGA_E_ARG
