"""The request engine: every request served in a mix (several jobs active,
admitted through the memory ledger as slots and memory free up) returns
exactly what its solo ParallelBeamModel.recon with the same seed returns."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from mbirtorch_amd import ParallelBeamModel
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    from mbirtorch_amd.serving import ReconRequest
    reqs = []
    cases = [(32, 4, 24), (48, 4, 36), (32, 8, 30), (64, 8, 20), (40, 4, 28)]
    for tag, (size, S, views) in enumerate(cases):
        angles = np.linspace(0, np.pi, views, endpoint=False)
        model = ParallelBeamModel((views, S, size), angles)
        rs = tuple(model.get_params('recon_shape'))
        vol = random_ellipsoids(rs, 4, seed=tag, axes_range=(2.0, size / 4))
        sino = add_noise(model.forward_project(vol), 0.01, seed=100 + tag)
        reqs.append(ReconRequest(sino, angles, max_iterations=4, seed=7 + tag,
                                 params=dict(p=1.2, q=2.0, T=1.0), tag=tag))
    return reqs


@pytest.mark.parametrize('max_active', [3, 1])
def test_mixed_results_equal_solo(max_active):
    from mbirtorch_amd.serving import ReconServer, run_plain
    print(f'[gpu] serving max_active={max_active}', flush=True)
    reqs = _requests()
    solo = {r.tag: r for r in run_plain(reqs)}
    got = {}
    with ReconServer(max_active=max_active) as srv:
        for r in reqs:
            srv.submit(r)
        while srv.pending():
            for res in srv.poll(timeout_ms=50):
                got[res.tag] = res
    assert sorted(got) == sorted(solo)
    for tag, res in got.items():
        ref = solo[tag]
        assert res.recon.shape == ref.recon.shape
        err = np.abs(res.recon - ref.recon).max()
        print(f'request {tag}: max |mix - solo| = {err:.3e}')
        assert np.array_equal(res.recon, ref.recon), tag
        np.testing.assert_allclose(res.fm_rmse, ref.fm_rmse, rtol=1e-12)
        np.testing.assert_allclose(res.alpha, ref.alpha, rtol=1e-6)


def test_ledger_queues_when_memory_is_short():
    from mbirtorch_amd import ParallelBeamModel
    from mbirtorch_amd.serving import ReconServer, request_device_bytes, MemoryPreflightError
    print('[gpu] serving ledger', flush=True)
    reqs = _requests()[:3]
    needs = []
    for r in reqs:
        m = ParallelBeamModel(r.sinogram.shape, r.angles)
        needs.append(request_device_bytes(r.sinogram.shape, m.get_params('recon_shape'),
                                          m.get_params('granularity'))['total'])
    lo = sorted(needs)
    cap = lo[0] + lo[1] - 1  # every request fits alone, no two fit together
    assert max(needs) <= cap
    with ReconServer(max_active=3) as srv:
        srv.ledger.capacity = cap
        for r in reqs:
            srv.submit(r)
        assert len(srv.jobs) == 1 and len(srv.queue) == 2
        done = []
        while srv.pending():
            done += srv.poll(timeout_ms=50)
            assert len(srv.jobs) <= 1
        assert sorted(d.tag for d in done) == [0, 1, 2]
        srv.ledger.capacity = 1
        with pytest.raises(MemoryPreflightError):
            srv.submit(_requests()[3])
