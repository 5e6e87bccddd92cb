"""The reference's hard-coded DLTC CC drivers (DLTC/ccd-t2.cc, ccsd-t1.cc,
ccsd-t2.cc) run layer by layer on the GPU, every tensor they assign (the
residual and each intermediate) against the einsum oracle of the same
statements, as whole-term GEMMs and as the reference's finest tile tasks."""
import numpy as np
import pytest

import oracle
from dtcf_amd import dltc_driver

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("tasks", [False, True], ids=["terms", "tile-tasks"])
@pytest.mark.parametrize("name", dltc_driver.DRIVERS)
def test_driver_vs_oracle(dev, name, tasks):
    O, V, wbO, wbV = 4, 8, 2, 4
    print(f"driver {name} O={O} V={V} tasks={tasks}", flush=True)
    d = dltc_driver.Driver(name, O, V, wbO, wbV, seed=3)
    arrays = {X: d.host(X) for X in d.t}
    d.run(tasks=tasks)
    ref = oracle.dltc_program_run(d.table, arrays)
    assigned = [X for X in d.t if dltc_driver.is_accumulator(X)]
    assert assigned
    for X in assigned:
        r = ref[X]
        err = np.max(np.abs(d.host(X) - r)) / max(np.max(np.abs(r)), 1e-300)
        assert err <= TOL, (name, X, err)
    # a second evaluation from zeroed accumulators gives the same result bitwise
    first = {X: d.host(X) for X in assigned}
    d.reset()
    d.run(tasks=tasks)
    for X in assigned:
        assert np.array_equal(d.host(X), first[X]), X
