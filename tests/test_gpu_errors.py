"""Failure handling of the C ABI, each condition forced from the host with
every kernel running to completion: a link cell over its 32-atom capacity at
start (host check), an atom over the neighbour-list capacity (the list build
stops writing at the capacity and raises a flag), an atom that jumps past the
neighbouring link cells between two link builds, and a non-finite momentum.
Each must come back as an SldError with its code and message."""
import numpy as np
import pytest

from helpers import consistent_state, fe_params
from spilady_amd import abi, configs

pytestmark = pytest.mark.gpu

SLD_ERR_CAPACITY, SLD_ERR_STATE = 3, 4


def _params(at, **kw):
    v = configs.fe_variables(300.0, 3, thermostats=True)
    v.update(kw)
    return abi.params_from_variables(v, at.box, at.natom)


def test_cell_over_capacity_at_start():
    from spilady_amd.engine import Engine, SldError
    at = configs.make_bcc_fe(10, 300.0, 3, a=1.5)  # 3^3 link cells of 74 atoms
    with pytest.raises(SldError) as e:
        Engine(_params(at), at, compute_initial_forces=False)
    assert e.value.code == SLD_ERR_CAPACITY and "link cell" in str(e.value)


def test_neighbour_list_over_capacity():
    from spilady_amd.engine import Engine, SldError
    at = configs.make_bcc_fe(10, 300.0, 3, a=2.2)  # 58 atoms within 4.1 A, 50 within 3.75 A
    with pytest.raises(SldError) as e:
        Engine(_params(at), at, compute_initial_forces=True)
    assert e.value.code == SLD_ERR_CAPACITY and "neighbours" in str(e.value)


def test_far_move_between_link_builds():
    from spilady_amd.engine import Engine, SldError
    v, p, at = fe_params(8, 300.0, 5, thermostats=True)
    st = consistent_state(p, at)
    e = Engine(p, st, compute_initial_forces=False)
    e.links()
    a = e.download()
    r = a["r"].copy()
    r[0, 0] += 2.0 * at.box[0] / 5.0  # two link cells along x
    e.upload({"r": r, "p": a["p"], "s": a["s"], "f": a["f"], "rho": a["rho"]})
    with pytest.raises(SldError) as err:
        e.links()
    assert err.value.code == SLD_ERR_STATE and "moved beyond" in str(err.value)
    e.close()


def test_non_finite_momentum_is_reported():
    from spilady_amd.engine import Engine, SldError
    v, p, at = fe_params(8, 300.0, 5, thermostats=True)
    st = consistent_state(p, at)
    e = Engine(p, st, compute_initial_forces=False)
    a = e.download()
    pp = a["p"].copy()
    pp[7, 1] = np.nan
    e.upload({"r": a["r"], "p": pp, "s": a["s"], "f": a["f"], "rho": a["rho"]})
    with pytest.raises(SldError) as err:
        e.step(0, 1)
    assert err.value.code == SLD_ERR_STATE and "non-finite" in str(err.value)
    e.close()
