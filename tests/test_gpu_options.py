"""DE options on the device (MI355X): low_froude 1 / 2 and sloped Manning
friction (the general instantiation of the flux and update kernels) against
the oracle's edge-once loop (equal to the reference's compiled C bit for bit,
tests/test_de_options.py): timestep sequences identical, states bit for bit.
Also with riverwalls, and the refusals of the paths that do not take them.
"""

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S
from tests.test_de_options import CASES

pytestmark = pytest.mark.gpu


def _compare(d, o):
    assert d.dt_history == o.dt_history
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


@pytest.mark.parametrize("builder,kw,ys,ft", CASES)
@pytest.mark.parametrize("low_froude,sloped", [(1, False), (2, False), (0, True), (2, True)])
def test_options_bitwise_vs_oracle(builder, kw, ys, ft, low_froude, sloped):
    print("case options %s lf=%d sloped=%s" % (builder.__name__, low_froude, sloped), flush=True)
    mesh, alg, q, bc = builder(**kw)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=False)), alg, q, bc)
    o.low_froude = low_froude
    o.use_sloped_mannings = sloped
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    d.set_low_froude(low_froude)
    d.set_sloped_mannings_function(sloped)
    for _ in d.evolve(yieldstep=ys, finaltime=ft):
        pass
    for _ in o.evolve(yieldstep=ys, finaltime=ft):
        pass
    _compare(d, o)


def test_options_with_riverwalls_host_sequenced():
    print("case options + riverwalls, host-sequenced", flush=True)
    mesh, alg, q, bc, walls, par = S.riverwall_levees(alg="DE1")
    eft, elev, rows, hp = S.riverwall_arrays(mesh, walls, par)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=False)), alg, q, bc)
    o.low_froude = 1
    o.use_sloped_mannings = True
    o.set_riverwalls(eft, elev, rows, hp)
    d = S.setup(Domain(mesh=mesh, reorder=False), alg, q, bc)
    d.set_low_froude(1)
    d.set_sloped_mannings_function(True)
    d.create_riverwalls(walls, par, verbose=False)
    d.device_timestep = False
    for _ in d.evolve(yieldstep=2.0, finaltime=8.0):
        pass
    for _ in o.evolve(yieldstep=2.0, finaltime=8.0):
        pass
    _compare(d, o)


def test_options_refused_on_mixed_and_gradients():
    print("case options refusals", flush=True)
    mesh, alg, q, bc = S.hills_de1(m=6, n=6)
    d = S.setup(Domain(mesh=mesh, precision="mixed"), alg, q, bc)
    d.set_low_froude(1)
    with pytest.raises(Exception, match="FP64"):
        for _ in d.evolve(yieldstep=1.0, finaltime=1.0):
            pass
    d.close()
    d = S.setup(Domain(mesh=mesh), "DE1", q, bc)
    d.set_sloped_mannings_function(True)
    with pytest.raises(Exception, match="flat friction"):
        d.adjoint_record(4, 0.05, [0, 1], 1)


def test_local_flux_updating_checks():
    """set_local_extrapolation_and_flux_updating keeps the reference's checks
    (shallow_water_domain.py:1193-1215); the library refuses the options its
    local-updating kernels do not carry instead of ignoring them."""
    from anuga_core_amd.domain import Domain
    from tests import scenarios as S
    mesh, _, q, bc = S.dam_break_de0(m=6, n=6)
    d = S.setup(Domain(mesh=mesh), "DE1", q, bc)
    with pytest.raises(Exception, match="only supported with euler"):
        d.set_local_extrapolation_and_flux_updating(nlevels=3)
    d.set_flow_algorithm("DE0")
    d.set_local_extrapolation_and_flux_updating(nlevels=3)
    for _ in d.evolve(yieldstep=0.1, finaltime=0.2):
        pass
    d.set_local_extrapolation_and_flux_updating(nlevels=0)     # 2**0 == 1: every step
    d.close()
    d = S.setup(Domain(mesh=mesh, precision="mixed"), "DE0", q, bc)
    with pytest.raises(Exception, match="FP64"):
        d.set_local_extrapolation_and_flux_updating(nlevels=3)
