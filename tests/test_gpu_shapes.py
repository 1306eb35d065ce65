"""Launch shapes, the autotuner (sld_set_shapes, sld_autotune) and the
graph modes (sld_set_graph_mode).

Every reduction inside the tunable kernels sums per wavefront, so the
trajectory is the same bit for bit under any valid shapes; the autotuner
times the candidates on the live state and restores it, so a run that is
tuned midway continues exactly as an untuned one."""
import numpy as np
import pytest

from helpers import consistent_state, fe_params
from spilady_amd import abi

pytestmark = pytest.mark.gpu

FIELDS = ("r", "p", "f", "s", "rho", "pe", "ke", "me", "me0", "stress", "heff_h", "heff_l")
# sweep_tail 0 launches every diffusion level on its own; the default (2048)
# and 8192 run the trailing levels of each sweep in one workgroup; spin_lanes
# 1 and 2 spread a cell's spin walk over fewer lanes than the default 4
OTHER_SHAPES = [
    dict(spin_block=32, sweep_block=256, force_block=128, nbuild_cells=1, temperature_block=64, sweep_tail=0,
         spin_lanes=1),
    dict(spin_block=128, sweep_block=128, force_block=64, nbuild_cells=4, temperature_block=128, sweep_tail=8192,
         spin_lanes=2),
]


def _obs_equal(a, b):
    return a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a)


def _full_physics(nu=8, T=600.0, seed=41):
    # all couplings except the renormalisation feedback, which drives the
    # electron temperature of so small a system negative within a few steps
    v, p, at = fe_params(nu, T, seed, thermostats=True, eltemp=True, localcolmot=True, renormalize=False)
    return p, consistent_state(p, at)


def _run(eng, nsteps, first=0):
    return eng.step(first, nsteps)


def test_trajectory_is_independent_of_launch_shapes():
    from spilady_amd.engine import Engine
    p, st = _full_physics()
    ref = Engine(p, st)
    obs_ref = _run(ref, 8)
    a_ref = ref.download()
    c_ref = ref.cells()
    ref.close()
    for shapes in OTHER_SHAPES:
        e = Engine(p, st)
        e.set_shapes(shapes)
        assert e.shapes() == shapes
        obs = _run(e, 8)
        for k in range(8):
            assert _obs_equal(obs[k], obs_ref[k]), (shapes, k)
        a = e.download()
        for f in FIELDS:
            np.testing.assert_array_equal(a[f], a_ref[f], err_msg=f"{shapes} {f}")
        c = e.cells()
        np.testing.assert_array_equal(c["te"], c_ref["te"])
        e.close()


def test_set_shapes_rejects_unsupported_shapes():
    from spilady_amd.engine import Engine, SldError
    p, st = _full_physics(6)
    e = Engine(p, st)
    with pytest.raises(SldError):
        e.set_shapes(dict(spin_block=96))
    with pytest.raises(SldError):
        e.set_shapes(dict(nbuild_cells=8))
    e.close()


def test_autotune_leaves_the_trajectory_unchanged():
    from spilady_amd.engine import Engine
    p, st = _full_physics()
    tuned = Engine(p, st)
    plain = Engine(p, st)
    o1 = _run(tuned, 3)
    o2 = _run(plain, 3)
    best, table = tuned.autotune(1)
    for name, cands in abi.SHAPE_CANDIDATES.items():
        assert best[name] in cands
        times = table[name]
        # every tunable kernel runs in a full-physics step
        assert all(t > 0.0 for t in times.values()), (name, times)
        assert times[best[name]] == min(times.values())
    assert tuned.shapes() == best
    o1 += _run(tuned, 5, 3)
    o2 += _run(plain, 5, 3)
    for k in range(8):
        assert _obs_equal(o1[k], o2[k]), k
    a1, a2 = tuned.download(), plain.download()
    for f in FIELDS:
        np.testing.assert_array_equal(a1[f], a2[f], err_msg=f)
    assert tuned.total_time == plain.total_time
    tuned.close()
    plain.close()


@pytest.mark.parametrize("print_every", [1, 3])
def test_graph_modes_give_identical_trajectories(print_every):
    """Eager launches, the replayed core_dTe graph and the replayed whole-step
    graphs (both cell-list parities, several replays, observables computed
    eagerly between replays) give the same trajectory bit for bit."""
    from spilady_amd.engine import Engine
    v, p, at = fe_params(8, 600.0, 43, thermostats=True, eltemp=True, localcolmot=True, print_every=print_every)
    st = consistent_state(p, at)
    runs = []
    for mode in (0, 1, 2):
        e = Engine(p, st)
        e.set_graph_mode(mode)
        obs = e.step(0, 5) + e.step(5, 7)
        runs.append((obs, e.download(), e.cells()))
        e.close()
    obs0, a0, c0 = runs[0]
    assert len(obs0) == 12 // print_every
    for obs, a, c in runs[1:]:
        assert len(obs) == len(obs0)
        for k in range(len(obs0)):
            assert _obs_equal(obs[k], obs0[k]), k
        for f in FIELDS:
            np.testing.assert_array_equal(a[f], a0[f], err_msg=f)
        np.testing.assert_array_equal(c["te"], c0["te"])


def test_graph_mode_rejects_unknown_modes():
    from spilady_amd.engine import Engine, SldError
    p, st = _full_physics(6)
    e = Engine(p, st)
    with pytest.raises(SldError):
        e.set_graph_mode(3)
    e.close()
