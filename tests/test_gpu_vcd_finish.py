"""The fused end of a subset step (mbir_vcd_finish) against the separate
launches it replaces, through the C-ABI: alpha from the partials, the error
update e -= alpha*A(delta), the recon update x[pix] += alpha*delta and its
per-slice statistics.  alpha, e and x are bit-identical (the same float32 and
FP64 operations); the statistics are FP64 sums in another fixed order, equal
to 1e-12 relative."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from mbirtorch_amd import _lib
    _lib.require_gpu()
    return _lib


def _case(rng, P, S, n_sino, n_dir, n_fwd, npix=4096):
    dev = 'cuda'
    pix = np.sort(rng.choice(npix, size=P, replace=False))
    pix[-1] = pix[0]  # a padded partition lists a pixel twice
    pix = np.sort(pix)
    d = rng.standard_normal((P, S)).astype(np.float32)
    dup = np.flatnonzero(pix[1:] == pix[:-1])
    d[dup + 1] = d[dup]  # a listed-twice pixel carries equal deltas
    dir_parts = rng.standard_normal(3 * n_dir) * 0.1  # (prior lin, prior quad, back lin)
    dir_parts[1::3] = np.abs(dir_parts[1::3]) + 1.0
    fwd_parts = rng.standard_normal(2 * n_fwd) * 0.1
    fwd_parts[1::2] = np.abs(fwd_parts[1::2]) + 1.0
    fwd_parts[0::2] += 0.5
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    return dict(
        pix=t(pix, torch.int64), delta=t(d, torch.float32),
        x=t(rng.standard_normal((npix, S)), torch.float32),
        e=t(rng.standard_normal(n_sino), torch.float32),
        adelta=t(rng.standard_normal(n_sino), torch.float32),
        dir_parts=t(dir_parts, torch.float64), fwd_parts=t(fwd_parts, torch.float64))


def _params(L, S):
    prm = L.VcdParams()
    prm.fm_constant = 0.7
    prm.max_alpha = 1.5
    prm.num_slices = S
    return prm


def _separate(L, c, prm, P, S, n_dir, n_fwd, alpha, alpha_sum, acc, scratch):
    lib, p, st = L.lib(), L.ptr, L.stream_ptr()
    L.check(lib.mbir_vcd_alpha(p(c['dir_parts']), n_dir, p(c['fwd_parts']), n_fwd,
                               ctypes.byref(prm), p(alpha), p(alpha_sum), st), 'alpha')
    L.check(lib.mbir_vcd_error_update(p(c['e']), p(c['adelta']), p(alpha), c['e'].numel(), st),
            'error update')
    L.check(lib.mbir_vcd_apply(p(c['x']), p(c['pix']), P, S, p(c['delta']), p(alpha), p(acc),
                               p(scratch), st), 'apply')


def _fused(L, c, prm, P, S, n_dir, n_fwd, alpha, alpha_sum, rows_buf):
    lib, p, st = L.lib(), L.ptr, L.stream_ptr()
    L.check(lib.mbir_vcd_finish(p(c['e']), p(c['adelta']), c['e'].numel(), p(c['x']),
                                p(c['pix']), P, S, p(c['delta']), p(c['dir_parts']), n_dir,
                                p(c['fwd_parts']), n_fwd, ctypes.byref(prm), p(alpha),
                                p(alpha_sum), p(rows_buf), st), 'finish')


@pytest.mark.parametrize('P,S,n_sino,n_dir,n_fwd', [
    (300, 16, 50_000, 19, 96),        # alpha formed in every workgroup
    (2048, 32, 4_000_000, 1024, 20000),  # past the re-read limit: one alpha workgroup
    (37, 5, 1_001, 1, 3),             # ragged: entry-wise update path
])
def test_finish_equals_separate_launches(P, S, n_sino, n_dir, n_fwd):
    L = _lib()
    lib = L.lib()
    prm = _params(L, S)
    rng = np.random.default_rng(P + S)
    rows = int(lib.mbir_vcd_block_stats_rows(P))
    scratch_n = int(lib.mbir_vcd_apply_scratch(P, S))
    assert rows * (S + 1) <= scratch_n
    c1 = _case(rng, P, S, n_sino, n_dir, n_fwd)
    c2 = {k: v.clone() for k, v in c1.items()}
    dev = 'cuda'
    a1 = torch.zeros(1, dtype=torch.float32, device=dev)
    a2 = torch.zeros(1, dtype=torch.float32, device=dev)
    s1 = torch.zeros(1, dtype=torch.float64, device=dev)
    s2 = torch.zeros(1, dtype=torch.float64, device=dev)
    acc1 = torch.zeros(S + 1, dtype=torch.float64, device=dev)
    acc2 = torch.zeros(S + 1, dtype=torch.float64, device=dev)
    scratch = torch.empty(scratch_n, dtype=torch.float64, device=dev)
    rows_buf = torch.zeros(scratch_n, dtype=torch.float64, device=dev)
    # Two steps into the same statistics rows, the second on a smaller subset.
    for step in range(2):
        Pk = P if step == 0 else max(P // 3, 2)
        _separate(L, c1, prm, Pk, S, n_dir, n_fwd, a1, s1, acc1, scratch)
        _fused(L, c2, prm, Pk, S, n_dir, n_fwd, a2, s2, rows_buf)
        torch.cuda.synchronize()
        assert a1.item() == a2.item()
    L.check(lib.mbir_vcd_block_stats_reduce(L.ptr(rows_buf), rows, S, L.ptr(acc2),
                                            L.stream_ptr()), 'reduce')
    torch.cuda.synchronize()
    assert s1.item() == s2.item()
    assert torch.equal(c1['e'], c2['e'])
    assert torch.equal(c1['x'], c2['x'])
    np.testing.assert_allclose(acc2.cpu().numpy(), acc1.cpu().numpy(), rtol=1e-12)
    assert torch.count_nonzero(rows_buf).item() == 0  # the reduction clears the rows


@pytest.mark.parametrize('n', [30_001, 30_000])  # scalar and four-entry update paths
def test_finish_bf16_and_read_alpha(n):
    """The bf16 error update rounds like mbir_vcd_error_update_bf16, and
    dir_parts = NULL with n_sino = 0 is a plain recon update with alpha read
    from the device."""
    L = _lib()
    lib, p, st = L.lib(), L.ptr, L.stream_ptr()
    from mbirtorch_amd.tomography_model import _to_bf16
    P, S = 500, 8
    prm = _params(L, S)
    c = _case(np.random.default_rng(7), P, S, n, 4, 40)
    alpha = torch.full((1,), 0.37, dtype=torch.float32, device='cuda')
    e2 = _to_bf16(c['e'])
    asum = torch.zeros(1, dtype=torch.float64, device='cuda')
    rows_buf = torch.zeros(int(lib.mbir_vcd_apply_scratch(P, S)), dtype=torch.float64,
                           device='cuda')
    L.check(lib.mbir_vcd_finish_bf16(p(e2), p(c['adelta']), n, p(c['x']), p(c['pix']), P, S,
                                     p(c['delta']), p(c['dir_parts']), 4, p(c['fwd_parts']), 40,
                                     ctypes.byref(prm), p(alpha), p(asum), p(rows_buf), st),
            'finish bf16')
    torch.cuda.synchronize()
    a = alpha.item()
    assert a != 0.37  # re-formed from the partials
    e1 = _to_bf16(c['e'])
    L.check(lib.mbir_vcd_error_update_bf16(p(e1), p(c['adelta']), p(alpha), n, st), 'upd')
    torch.cuda.synchronize()
    assert torch.equal(e1.view(torch.int16), e2.view(torch.int16))
    # Read mode: alpha fixed, no error update.
    x0 = c['x'].clone()
    alpha.fill_(0.25)
    L.check(lib.mbir_vcd_finish(None, None, 0, p(c['x']), p(c['pix']), P, S, p(c['delta']),
                                None, 0, None, 0, None, p(alpha), None, p(rows_buf), st),
            'finish read')
    torch.cuda.synchronize()
    assert alpha.item() == 0.25 and a != 0.25
    ref = x0.clone()
    ref.index_add_(0, c['pix'], 0.25 * c['delta'])
    torch.testing.assert_close(c['x'], ref, rtol=0, atol=1e-6)


def test_linear_term_from_back_projection():
    """<W e, A delta> from the line search with the error sinogram equals
    <A^T W e, delta> from the weighted back projection (the direction
    partials' third value), and the line search without the error sinogram
    forms the same quadratic term with a zero linear term."""
    L = _lib()
    lib, p, st = L.lib(), L.ptr, L.stream_ptr()
    import mbirtorch_amd as mb
    V, R, C = 60, 8, 48
    model = mb.ParallelBeamModel((V, R, C), np.linspace(0, np.pi, V, endpoint=False))
    ops = model.projector()
    rs = tuple(model.get_params('recon_shape'))
    rng = np.random.default_rng(3)
    np.random.seed(3)
    sub = torch.as_tensor(mb.gen_pixel_partition(rs, 8)[0].astype(np.int64), device='cuda')
    P = int(sub.shape[0])
    d = torch.as_tensor(rng.standard_normal((P, R)), dtype=torch.float32, device='cuda')
    e = torch.as_tensor(rng.standard_normal((V, C, R)), dtype=torch.float32, device='cuda')
    w = torch.as_tensor(rng.random((V, C, R)) + 0.5, dtype=torch.float32, device='cuda')
    n = ops.fwd_partials_count()
    pa = torch.empty(2 * n, dtype=torch.float64, device='cuda')
    pb = torch.empty(2 * n, dtype=torch.float64, device='cuda')
    ops.fwd_linesearch(d, sub, e, w, pa)
    ops.fwd_linesearch(d, sub, None, w, pb)
    bt = ops.back(e, sub, weights=w)  # A^T W e at the subset, (P, R)
    torch.cuda.synchronize()
    a, b = pa.view(n, 2).cpu().numpy(), pb.view(n, 2).cpu().numpy()
    assert np.array_equal(a[:, 1], b[:, 1]) and not b[:, 0].any()
    back_lin = float((bt.double() * d.double()).sum())
    assert abs(a[:, 0].sum() - back_lin) <= 1e-5 * np.abs(a[:, 0]).sum()
