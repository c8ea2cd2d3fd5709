"""The continuous-batching request loop (serving.py) on the GPU: a short
Poisson stream of uneven requests (vortex, Taylor-Green with and without the
dynamic Smagorinsky model) runs through the real loop, joining and leaving
the batch while it runs; every request's result equals its solo run bit for
bit, in both stage modes."""
import pytest

import chimera_comp_les_amd as cl
from chimera_comp_les_amd import serving as sv

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,graph", [(cl.BATCH_PERSISTENT, True), (cl.BATCH_LAUNCHES, True),
                                        (cl.BATCH_LAUNCHES, False)])
def test_request_stream_equals_solo_runs(mode, graph):
    reqs = sv.make_requests(n=14, rate=400.0, steps=7, seed=11, sizes=(16, 20, 24, 32), les_share=0.3)
    assert {r.kind for r in reqs} == {"vortex", "tgv", "tgv-les"}
    pay = sv.payloads(reqs)
    seq = sv.run_sequential(reqs, pay)
    bat = sv.run_batched(reqs, pay, budget=1 << 36, tile=7, chunk=3, policy="aging", cap_cells=40000, age_s=0.05,
                         mode=mode, graph=graph)
    assert bat["max_concurrency"] > 1
    for r in reqs:
        assert bat["digests"][r.rid] == seq["digests"][r.rid], r
