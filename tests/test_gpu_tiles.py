"""The tile height of the stage kernel is a work decomposition only: for every
height the states are bit-identical and the accepted / rejected timestep
counts equal (the FP64 volume is summed per tile first, so it agrees to
rounding).  Heights from a single row (a march of three rows, odd in the
FP32 form's two-row loop) up to above the grid's rows, with a ragged last
tile, on the storm catchment (drifting storm, Voronoi roughness, physics
flags) in both precisions."""
import numpy as np
import pytest

from mfcm_amd import scenarios as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-tiles]", *a, flush=True)


def _run(sc, prec, rows):
    from mfcm_amd.mfsv import Catchment
    c = Catchment(sc, precision=prec)
    c.set_tile_rows(rows)
    c.run_begin()
    st = c.run()
    res = (c.state(), st["accepted"], st["rejected"], st["volume"])
    c.close()
    return res


@pytest.mark.parametrize("prec_name", ["f32", "f64"])
def test_tile_height_does_not_change_results(gpu_lib, prec_name):
    from mfcm_amd import mfsv
    prec = mfsv.F32 if prec_name == "f32" else mfsv.F64
    sc = S.storm_catchment(seed=1, n=130, dx=2.0, t_end=120.0, prewet=0.002)
    _say(prec_name, "storm catchment 130x130, rows 18 (reference)")
    state0, acc0, rej0, vol0 = _run(sc, prec, 18)
    assert acc0 > 0
    for rows in (1, 2, 3, 4, 6, 8, 13, 64):
        _say(prec_name, "rows", rows)
        state, acc, rej, vol = _run(sc, prec, rows)
        assert (acc, rej) == (acc0, rej0), rows
        for x, y in zip(state, state0):
            np.testing.assert_array_equal(x, y)
        assert abs(vol - vol0) <= 1e-12 * abs(vol0), (rows, vol, vol0)
