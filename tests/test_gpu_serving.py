"""Continuous-batching server: a stream of small flood scenarios, admitted
under an HBM budget that holds only some of them at once; each result equals
the scenario run alone (bit-identical depth field and timestep counts)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-serving]", *a, flush=True)


def test_server_results_equal_alone(gpu_lib):
    from mfcm_amd import mfsv
    from mfcm_amd.serving import FloodServer, Request, footprint_bytes, summarize

    prec = mfsv.F32
    reqs = [Request(k, 500 + k, arrival=0.05 * k, side_lo=40, side_hi=160, horizon=90.0)
            for k in range(6)]
    scs = [r.scenario() for r in reqs]
    _say("sides", [(s.nx, s.ny) for s in scs])
    budget = sum(sorted(footprint_bytes(s.nx, s.ny, prec) for s in scs)[-3:])
    srv = FloodServer(prec, hbm_budget=budget, chunk=16)
    done = srv.serve(reqs, wall_limit=100.0, keep_fields=True)
    srv.close()
    assert len(done) == len(reqs) and srv.backlog == 0
    peak = 0.0
    for r in done:
        sc = r.scenario()
        c = mfsv.Catchment(sc, precision=prec)
        c.run_begin()
        st = c.run()
        h = c.download("h")
        c.close()
        assert (r.result["accepted"], r.result["rejected"]) == (st["accepted"], st["rejected"])
        np.testing.assert_array_equal(r.result["h"], h)
        assert r.result["time_to_result"] >= 0.0
        peak = max(peak, r.result["time_to_result"])
    s = summarize(done, srv.elapsed)
    _say("summary", {k: v for k, v in s.items()})
    assert s["completed"] == 6 and s["p99_s"] <= peak + 1e-12


def test_admission_respects_budget(gpu_lib):
    from mfcm_amd import mfsv
    from mfcm_amd.serving import FloodServer, Request, footprint_bytes

    prec = mfsv.F32
    reqs = [Request(k, 900 + k, arrival=0.0, side_lo=48, side_hi=64, horizon=30.0)
            for k in range(4)]
    one = max(footprint_bytes(r.scenario().nx, r.scenario().ny, prec) for r in reqs)
    srv = FloodServer(prec, hbm_budget=one * 1.5, chunk=8)
    seen = []
    orig = srv._admit

    def admit(req, now):
        ok = orig(req, now)
        seen.append(len(srv.active))
        return ok
    srv._admit = admit
    done = srv.serve(reqs, wall_limit=60.0)
    srv.close()
    assert len(done) == 4
    assert max(seen) == 1          # never two resident at once under 1.5x one footprint
