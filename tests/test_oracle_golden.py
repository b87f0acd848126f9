"""The oracle against the reference's own golden vectors.

The reference keeps the receiver traces of its composite test flow
(tests/test1-cmp-eci-ii/eci/SConstruct with par1.py / par3.py: 2 grids at
ratio 2, 385x385x73 + 193x193x60 nodes, 1172 steps, free surface, vx force
Ricker) in its SVN pristine store; tests/golden/reference_traces.npz holds
them, decoded by tests/golden/decode_reference_traces.py.

Over the full 1172 steps the oracle reproduces the ECI traces of tests
1/2/3 to 1.35e-7 / 1.36e-7 / 1.0e-7 relative L2 and the II traces to
1.35e-7 / 1.36e-7 / 6.4e-8, against 1.38e-2 / 7.2e-2 / 0.83 between the
modes; that is the size of the float32 rounding of the golden samples.  The
grid steps are those the input files' headers hold: sfspike writes d1..d3
with %g, so the reference reads h = 4.16667 and 8.33333 m (the flow script's
4.166667 and 8.333333 leave a 1.4e-5 residual, a delay that grows with
offset).  The full runs take minutes on 8 cores and are marked slow.  The
default suite checks prefixes.

The II flows (tests/test1-cmp-eci-ii/ii, imode=0) are matched with the
x-end vz ghost edge left unwritten (ii_xend 0), as the committed version of
update_ghost_vz.cpp:98 does; the working copy's form of that line drives the
test3 traces 0.16 away by step 1000.
"""
import os

import numpy as np
import pytest

from ecsg_amd.model import reference_flow
from helpers import rel
from oracle_lib import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "reference_traces.npz"))


def test_golden_file_layout():
    for t in (1, 2, 3):
        for mode in ("eci", "ii"):
            a = GOLD["rec_%s_test%d" % (mode, t)]
            assert a.shape == (4, 1172) and a.dtype == np.float32
            assert abs(float(GOLD["rec_%s_test%d_d1" % (mode, t)]) - 0.589378) < 1e-9
    geo = GOLD["geo_eci_test1"]
    assert np.allclose(geo[:, 0], [875.0, 958.3333, 1033.3334, 1116.6666], atol=1e-3)


@pytest.mark.parametrize("test,steps,tol", [(1, 150, 2e-7), (3, 200, 2e-7)])
def test_oracle_matches_golden_prefix(test, steps, tol):
    m = reference_flow(test)
    o = Oracle(m, eci_mode=0, check_energy=False)
    o.run(steps)
    tr = o.traces()
    g = GOLD["rec_eci_test%d" % test][:, :steps]
    err = rel(tr, g)
    assert err <= tol, err
    o.close()


@pytest.mark.slow
@pytest.mark.parametrize("test", [1, 2, 3])
def test_oracle_matches_golden_full(test):
    m = reference_flow(test)
    o = Oracle(m, eci_mode=0, check_energy=True)
    o.run(m.nsteps)
    tr = o.traces()
    g = GOLD["rec_eci_test%d" % test]
    err = rel(tr, g)
    err_ii = rel(tr, GOLD["rec_ii_test%d" % test])
    assert err <= 3e-7 and err_ii >= 1000 * err
    e = o.energy()
    assert abs(e[-1] - e[701]) / e[701] <= 1e-6


def test_oracle_ii_matches_golden_prefix():
    """II interpolation (imode 0) against the II flow's golden traces.  With
    the source next to the interface (par3.py) the II and ECI goldens are
    0.25 apart after 200 steps, so this prefix tells the modes apart."""
    m = reference_flow(3, eci=False)
    assert m.imode == 0
    o = Oracle(m, eci_mode=0, check_energy=False)
    o.run(200)
    tr = o.traces()
    err = rel(tr, GOLD["rec_ii_test3"][:, :200])
    err_eci = rel(tr, GOLD["rec_eci_test3"][:, :200])
    assert err <= 2e-7, err
    assert err_eci >= 1000 * err
    o.close()


@pytest.mark.slow
@pytest.mark.parametrize("test", [1, 2, 3])
def test_oracle_ii_matches_golden_full(test):
    m = reference_flow(test, eci=False)
    o = Oracle(m, eci_mode=0, check_energy=False)
    o.run(m.nsteps)
    tr = o.traces()
    err = rel(tr, GOLD["rec_ii_test%d" % test])
    err_eci = rel(tr, GOLD["rec_eci_test%d" % test])
    assert err <= 3e-7 and err_eci >= 1000 * err, (err, err_eci)
    o.close()
