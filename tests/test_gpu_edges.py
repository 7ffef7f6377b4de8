"""Edge cases and error paths of the device layer: empty queries, arguments
out of range, and the conditions the reference turns into print-and-exit
(cell_list.c:126-129 cells narrower than a diameter, cell_list.c:229-232 a
cell over max_part).  Every error is a status code; the context stays usable
after it."""
import numpy as np
import pytest

from hsmc_amd import gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BAD_ARG, CELL_OVERFLOW, CELL_TOO_SMALL, NOT_SUPPORTED = -5, -1, -2, -8


def _code(fn):
    with pytest.raises(gpu.HsmcError) as e:
        fn()
    return e.value.code


@pytest.fixture(scope="module")
def sys500():
    o = O.System(2, 5, rho=0.5, neigh_dr=1.0, max_part=10, seed=1)
    conf = o.config()
    L = o.box()["lx"]
    ctx = gpu.Context([dict(n=len(conf), box=(L, L, L), dr_max=0.1, seed=3)])
    ctx.upload(conf)
    yield ctx, conf, L
    ctx.close()


def test_box_limits():
    # a 2-cell sweep grid needs an edge above 2 diameters, and cells narrower
    # than 1.5 (edges in [3, 4) would give cells of 1.5 to 2)
    assert _code(lambda: gpu.Context([dict(n=4, box=(1.9, 5.0, 5.0))])) == BAD_ARG
    assert _code(lambda: gpu.Context([dict(n=8, box=(3.5, 6.0, 6.0))])) == NOT_SUPPORTED
    ctx = gpu.Context([dict(n=8, box=(2.9, 4.1, 6.2))])   # cells 1.45, 1.025, 1.033
    assert ctx.grid() == (2, 4, 6)
    ctx.close()


def test_empty_queries(sys500):
    ctx, conf, L = sys500
    assert len(ctx.overlap_batch(np.zeros(0, np.int32), np.zeros((0, 3)))) == 0
    assert len(ctx.widom_points(np.zeros((0, 3)))) == 0
    assert len(ctx.pair_dist(np.zeros(0, np.int32), np.zeros(0, np.int32))) == 0
    assert ctx.widom_uniform(1, 0, 0, 0) == 0
    ctx.sweeps(0)
    counts = ctx.npt_sweeps(0, 3.0, 0.01)
    assert counts.sum() == 0
    assert np.array_equal(ctx.download(), conf)


def test_arguments_out_of_range(sys500):
    ctx, conf, L = sys500
    bad = conf.copy()
    bad[7, 2] = L + 0.5
    assert _code(lambda: ctx.upload(bad)) == BAD_ARG
    assert _code(lambda: ctx.rdf_counts(0.01, 400, L / 2 + 0.5)) == BAD_ARG   # beyond minimum image
    assert _code(lambda: ctx.any_overlap_scaled(np.full(65, 0.99))) == BAD_ARG   # at most 64 factors
    assert _code(lambda: ctx.set_tuning(48, 2, 0)) == BAD_ARG
    assert _code(lambda: ctx.set_tuning(32, 3, 0)) == BAD_ARG   # cells per workgroup: 1, 2, 4 or 8
    assert _code(lambda: ctx.set_tuning(64, 8, 0)) == BAD_ARG   # at most 256 lanes
    assert _code(lambda: ctx.npt_sweeps(1, -1.0, 0.01)) == BAD_ARG
    assert _code(lambda: ctx.npt_sweeps(1, 3.0, 0.0)) == BAD_ARG
    assert _code(lambda: ctx.overlap_batch(np.array([len(conf)], np.int32), np.zeros((1, 3)))) == BAD_ARG
    assert _code(lambda: ctx.box(1)) == BAD_ARG   # one replica only
    net = gpu.Mlp(ctx, seed=1)
    assert _code(lambda: net.widom(0.25, 1.0, 1, 0, 1000)) == BAD_ARG   # mix must be below 1
    net.close()
    # nothing above changed the configuration
    assert np.array_equal(ctx.download(), conf)


def test_reference_exit_conditions(sys500):
    ctx, conf, L = sys500
    # cells narrower than a diameter (cell_list.c:126-129)
    assert _code(lambda: ctx.cell_list((11, 10, 10), 10)) == CELL_TOO_SMALL
    # a cell with more occupants than max_part - 1 (cell_list.c:229-232):
    # 500 spheres in 27 cells of side 3.33
    assert _code(lambda: ctx.cell_list((3, 3, 3), 10)) == CELL_OVERFLOW
    # the context is usable afterwards
    pc, nm = ctx.cell_list((10, 10, 10), 10)
    assert pc[:, 0].sum() == len(conf)
    ctx.sweeps(5)
    ctx.sync()
    assert ctx.any_overlap_scaled([1.0])[0] == 1
