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


def test_batched_blocks_same_as_alone():
    """Requests of one grid as the blocks of one batch solve (own freestream,
    per-block norms and forces): every block's state is bit-identical to the
    same request solved alone, its norms agree to rounding (per-block sums
    in another order) and its force coefficients are bit-identical."""
    from cosa_amd.solver import Solve
    reqs = [serve.Request(rid=i, grid="small", mach=m, alpha_deg=a) for i, (m, a) in
            enumerate([(0.5, 1.0), (0.35, -1.5), (0.45, 3.0)])]
    c = serve.batch_case(reqs[0], 3)
    sv = Solve(c, device=0, q=c.freestream())
    try:
        sv.set_block_rms(True)
        for i, r in enumerate(reqs):
            sv.set_block_freestream(i + 1, r.mach, r.alpha_deg)
        for nl in range(1, c.config["nlevel"] + 1):   # every level starts from the request's freestream
            q = sv.download("q", nl)
            for i, r in enumerate(reqs):
                q[serve.block_slice(c, i, nl)] = r.make_case().freestream(nl)
            sv.upload("q", q, nl)
        rms = []
        for mgit in range(1, 6):
            sv.cycle_enqueue(mgit)
            sv.cycle_wait()
            rms.append(sv.block_rms())
        fb = sv.force_blocks(nsurface=1, include_inviscid=True)
        qb = sv.download("q")
    finally:
        sv.close()
    for i, r in enumerate(reqs):
        one = r.make_case()
        a = Solve(one, device=0, q=one.freestream())
        try:
            for mgit in range(1, 6):
                a.cycle_enqueue(mgit)
                a.cycle_wait()
            assert np.array_equal(a.download("q"), qb[serve.block_slice(c, i)]), f"block {i}"
            for m in range(5):
                np.testing.assert_allclose(np.array(rms[m][i][:6]), np.array(a.history[m][:6]), rtol=1e-12, atol=0)
            fa = a.force(nsurface=1, include_inviscid=True)
            assert np.array_equal(fa, fb[i]), (fa[0, 0, :2], fb[i, 0, 0, :2])
            print(f"  block {i}: M {r.mach} alpha {r.alpha_deg}: CL_p {fb[i, 0, 0, 0]:.6f} CD_p {fb[i, 0, 0, 1]:.6f}, "
                  f"resrms(1) {rms[-1][i][0]:.3e}", flush=True)
        finally:
            a.close()


def test_batch_server_retires_every_request():
    reqs = [serve.Request(rid=i, grid=g, mach=0.3 + 0.05 * i, alpha_deg=0.5 * i, arrival=0.01 * i, drop=0.5,
                          max_cycles=30) for i, g in enumerate(["small", "small", "medium", "small", "medium"])]
    res = serve.BatchServer(keep_state=True).run(reqs)
    assert sorted(r.rid for r in res) == [0, 1, 2, 3, 4]
    for r in res:
        assert r.cycles >= 1 and np.all(np.isfinite(r.forces)) and r.q is not None
    summary = serve.summarize(res, max(r.t_done for r in res), serve.cell_updates(res))
    print(f"batch server: {summary}", flush=True)


def test_rans_request_in_mix_same_as_alone():
    """A steady k-omega RANS request (viscous wall, Wilcox k-omega, the
    serving mix's RANS share) served beside an Euler and an HB request:
    bit-identical to the same request served alone, and retired by the
    stall rule or its drop like any other."""
    reqs = [serve.Request(rid=0, grid="small", mach=0.6, alpha_deg=4.0, rans=True, arrival=0.0, drop=1e-2,
                          max_cycles=30, stall_window=20),
            serve.Request(rid=1, grid="small", mach=0.3, alpha_deg=-2.0, arrival=0.0, drop=1e-2, max_cycles=30),
            serve.Request(rid=2, grid="small", mach=0.7, alpha_deg=6.0, nharms=2, arrival=0.01, drop=1e-2,
                          max_cycles=30)]
    mixed = {r.rid: r for r in serve.Server(keep_state=True).run(reqs)}
    alone = serve.Server(keep_state=True).run([reqs[0]])[0]
    r = mixed[0]
    print(f"RANS request: {r.cycles} cycles, drop reached {r.drop_reached:.2e}, converged {r.converged}, "
          f"stalled {r.stalled}, CL {r.forces[0, 0] + r.forces[0, 6]:.6f}", flush=True)
    assert r.q.size == alone.q.size and np.array_equal(r.q, alone.q)
    assert r.resrms == alone.resrms and r.cycles == alone.cycles
    assert np.array_equal(r.forces, alone.forces)
    assert np.isfinite(r.resrms).all() and r.drop_reached < 1.0
