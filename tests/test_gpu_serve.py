"""Serving loop (cosa_amd.serve): requests admitted while others are on
the GPU, their cycles interleaved on separate streams, retired at their
residual drop.  Each result -- final state, residual history, forces -- must
be bit-identical to the same request served alone: the solves share the
device but no data, and each one's launch sequence is the same in the mix.
"""
import numpy as np
import pytest

from cosa_amd import serve

pytestmark = pytest.mark.gpu


def _requests(precision):
    return [serve.Request(rid=0, grid="small", mach=0.5, alpha_deg=1.0, arrival=0.0, drop=1e-2, max_cycles=40,
                          precision=precision),
            serve.Request(rid=1, grid="small", mach=0.4, alpha_deg=2.5, nharms=1, arrival=0.0, drop=1e-2,
                          max_cycles=40, precision=precision),
            serve.Request(rid=2, grid="medium", mach=0.6, alpha_deg=0.5, arrival=0.02, drop=1e-2, max_cycles=40,
                          precision=precision),
            serve.Request(rid=3, grid="small", mach=0.35, alpha_deg=3.0, arrival=0.05, drop=5e-3, max_cycles=40,
                          precision=precision)]


@pytest.mark.parametrize("precision", [0, 2])
def test_each_result_same_alone_and_in_the_mix(precision):
    reqs = _requests(precision)
    mixed = serve.Server(keep_state=True).run(reqs)
    assert sorted(r.rid for r in mixed) == [0, 1, 2, 3]
    summary = serve.summarize(mixed, max(r.t_done for r in mixed), serve.cell_updates(mixed))
    print(f"precision mode {precision}: {summary}", flush=True)
    # the mix really overlapped: some request was admitted before another retired
    spans = sorted((r.t_admit, r.t_done) for r in mixed)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    for r in mixed:
        alone = serve.Server(keep_state=True).run([reqs[r.rid]])[0]
        print(f"  request {r.rid}: {r.cycles} cycles, resrms(1) {r.resrms[0]:.3e} -> {r.resrms[-1]:.3e}, "
              f"CL_p {r.forces[0, 0]:.6f}, latency {r.latency:.3f} s (alone {alone.latency:.3f} s)", flush=True)
        assert r.cycles == alone.cycles and r.converged == alone.converged
        assert np.array_equal(r.q, alone.q)
        assert r.resrms == alone.resrms
        assert np.array_equal(r.forces, alone.forces)
