"""Batched catchments: every member's result is bit-identical alone and in
the mix (same accepted/rejected timestep counts, same fields), also when a
member joins a batch that is already running."""
import numpy as np
import pytest

from mfcm_amd import scenarios as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-batch]", *a, flush=True)


def _members(n, t_end=120.0):
    out = []
    for k in range(n):
        sc = S.ensemble_member(100 + k, side_lo=40, side_hi=180, t_end=t_end)
        out.append(sc)
    return out


def _alone(sc, prec):
    from mfcm_amd.mfsv import Catchment
    c = Catchment(sc, precision=prec)
    c.run_begin()
    st = c.run()
    res = (c.state(), st["accepted"], st["rejected"])
    c.close()
    return res


@pytest.mark.parametrize("prec_name", ["f32", "f64"])
def test_batch_equals_alone(gpu_lib, prec_name):
    from mfcm_amd import mfsv
    from mfcm_amd.batch import Batch
    prec = mfsv.F32 if prec_name == "f32" else mfsv.F64
    scs = _members(5)
    _say(prec_name, "members", [(s.nx, s.ny, round(s.dx, 2)) for s in scs])
    ref = [_alone(sc, prec) for sc in scs]
    b = Batch(prec)
    cats = []
    for sc in scs:
        c = mfsv.Catchment(sc, precision=prec)
        c.run_begin()
        b.add(c)
        cats.append(c)
    left = b.run()
    assert left == 0
    for c, (state, acc, rej) in zip(cats, ref):
        st = c.stats()
        assert st["done"] == 1
        assert (st["accepted"], st["rejected"]) == (acc, rej)
        for x, y in zip(c.state(), state):
            np.testing.assert_array_equal(x, y)
    b.close()
    for c in cats:
        c.close()


def test_member_joining_a_running_batch(gpu_lib):
    from mfcm_amd import mfsv
    from mfcm_amd.batch import Batch
    prec = mfsv.F32
    scs = _members(3, t_end=90.0)
    ref = [_alone(sc, prec) for sc in scs]
    b = Batch(prec)
    cats = [mfsv.Catchment(sc, precision=prec) for sc in scs]
    for c in cats[:2]:
        c.run_begin()
        b.add(c)
    b.run(40)                      # the first two are part-way
    cats[2].run_begin()
    b.add(cats[2])                 # continuous batching: a new arrival joins
    b.run()
    for c, (state, acc, rej) in zip(cats, ref):
        st = c.stats()
        assert (st["accepted"], st["rejected"]) == (acc, rej)
        for x, y in zip(c.state(), state):
            np.testing.assert_array_equal(x, y)
    b.close()


@pytest.mark.parametrize("prec_name", ["f32", "f64"])
def test_batch_with_further_sources_equals_alone(gpu_lib, prec_name):
    """Members with infiltration, energy loss, Smagorinsky eddy viscosity or
    a coded-out (NaN) bed cell share a launch with plain members: the batch
    takes its general kernel, each member its own tile code."""
    from mfcm_amd import mfsv
    from mfcm_amd.batch import Batch
    prec = mfsv.F32 if prec_name == "f32" else mfsv.F64
    scs = _members(4, t_end=90.0)
    rng = np.random.default_rng(5)
    s1 = scs[1]
    s1.infiltration = (rng.uniform(0, 1e-5, (s1.ny, s1.nx)), np.full((s1.ny, s1.nx), 2e-3))
    s1.energy_loss = (0.01, 0.02)
    scs[2].eddy = {"smagorinsky": (0.05, 0.5)}
    zb = scs[3].zb.copy()
    zb[scs[3].ny // 2, scs[3].nx // 2] = np.nan
    scs[3].zb = zb
    _say(prec_name, "members", [(s.nx, s.ny) for s in scs])
    ref = [_alone(sc, prec) for sc in scs]
    b = Batch(prec)
    cats = []
    for sc in scs:
        c = mfsv.Catchment(sc, precision=prec)
        c.run_begin()
        b.add(c)
        cats.append(c)
    assert b.run() == 0
    for c, (state, acc, rej) in zip(cats, ref):
        st = c.stats()
        assert (st["accepted"], st["rejected"]) == (acc, rej)
        for x, y in zip(c.state(), state):
            np.testing.assert_array_equal(x, y)
    b.close()
    for c in cats:
        c.close()
