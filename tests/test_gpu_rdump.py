"""Restart records (the rdump files of restart.c) packed and unpacked on the GPU,
through the C ABI, against the oracle's restatement of rdump_block_write /
rdump_block_read (restart.c:52-67, :178-269).

Workloads: the torus state of the 64x32x32 grid in 4x4x4 blocks of 16x8x8
(a = 0.9375, Gamma = 13/9, FM torus r_in = 6, r_max = 12, seeded +-4%
perturbation of u, ghosts filled), and the three-level set of uneven block
sizes of the multi-level tests (each block with its own level's metric).

Bars: every record bit-exact (copies and one product per face value); every
read bit-exact against the oracle, for records of the same block size and of
blocks (2,2,2) and (1,2,3) times finer (the device sums in the reference's
visiting order); the whole-grid device write equal to the per-block writes;
a write -> files -> read -> ghost fill round trip restores the primitives bit
for bit and the face field to within two roundings ((x gdet) / gdet, and the
theta-fraction ratio of the torus grid); a run resumed from the files in a
fresh context follows the uninterrupted run to 1e-9 relative after two more
steps; the dump records (9 floats per interior cell) bit-exact.
"""
import numpy as np
import pytest
import torch

import hamr_gamma_wind_amd as H

pytestmark = pytest.mark.gpu

NG, NPR = 3, 9
N2 = 32


def _gdet(ctx, geo, D1, D2):
    return np.ascontiguousarray(ctx.download_metric(geo)["gdet"][:, :D1 * D2])


@pytest.fixture(scope="module")
def uni():
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    p0 = np.ascontiguousarray(ctx.get(H.HAMR_P))
    ps0 = np.ascontiguousarray(ctx.get(H.HAMR_PS))
    D = ctx.D
    g = [_gdet(ctx, int(ctx.block_row(n)[0]), D[0], D[1]) for n in range(ctx.nblk)]
    yield ctx, p0, ps0, g
    ctx.close()


def _flat(a):
    return np.ascontiguousarray(a).reshape(a.shape[0], -1)


def test_block_write_matches_oracle(uni, oracle):
    ctx, p0, ps0, g = uni
    bs = [d - 2 * NG for d in ctx.D]
    for n in range(ctx.nblk):
        rec = ctx.rdump_block_write(n)
        want = oracle.rdump_write(bs, _flat(p0[n]), _flat(ps0[n]), g[n])
        assert np.array_equal(rec, want), n
    print("records of %d blocks bit-exact" % ctx.nblk, flush=True)


def test_write_dev_equals_block_writes(uni):
    ctx, p0, ps0, g = uni
    off = ctx.rdump_offsets()
    buf = torch.zeros(int(off[-1]), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.rdump_write_dev(buf.data_ptr())
    got = buf.cpu().numpy()
    for n in range(ctx.nblk):
        assert np.array_equal(got[off[n]:off[n + 1]], ctx.rdump_block_write(n)), n


def _perturbed(p0, ps0, seed):
    rng = np.random.default_rng(seed)
    return p0 * rng.uniform(0.5, 1.5, p0.shape), ps0 + rng.standard_normal(ps0.shape)


def test_block_read_matches_oracle(uni, oracle):
    ctx, p0, ps0, g = uni
    bs = [d - 2 * NG for d in ctx.D]
    recs = [ctx.rdump_block_write(n) for n in range(ctx.nblk)]
    pb, psb = _perturbed(p0, ps0, 3)
    ctx.set(H.HAMR_P, pb)
    ctx.set(H.HAMR_PS, psb)
    for n in range(ctx.nblk):
        ctx.rdump_block_read(n, recs[n])
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    for n in range(ctx.nblk):
        wp, wps = oracle.rdump_read(bs, (1, 1, 1), 1, N2, recs[n], g[n], _flat(pb[n]), _flat(psb[n]))
        assert np.array_equal(_flat(gp[n]), wp) and np.array_equal(_flat(gps[n]), wps), n
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)


@pytest.mark.parametrize("red", [(2, 2, 2), (1, 2, 3)])
def test_reduced_read_matches_oracle(uni, oracle, red):
    ctx, p0, ps0, g = uni
    bs = [d - 2 * NG for d in ctx.D]
    rng = np.random.default_rng(sum(red))
    for n in (0, 5, ctx.nblk - 1):
        size = ctx.rdump_size(n, red)
        assert size == (NPR + 4) * np.prod([b * r + 2 * NG for b, r in zip(bs, red)])
        rec = rng.standard_normal(size)
        ctx.rdump_block_read(n, rec, red)
        gp, gps = ctx.get(H.HAMR_P)[n], ctx.get(H.HAMR_PS)[n]
        wp, wps = oracle.rdump_read(bs, red, 1, N2, rec, g[n], _flat(p0[n]), _flat(ps0[n]))
        assert np.array_equal(_flat(gp), wp) and np.array_equal(_flat(gps), wps), (n, red)
        ctx.set(H.HAMR_P, p0)
        ctx.set(H.HAMR_PS, ps0)
    print("reads of records %s times finer bit-exact" % (red,), flush=True)


def test_read_dev_matches_oracle(uni, oracle):
    ctx, p0, ps0, g = uni
    bs = [d - 2 * NG for d in ctx.D]
    off = ctx.rdump_offsets()
    buf = torch.zeros(int(off[-1]), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.rdump_write_dev(buf.data_ptr())
    rec = buf.cpu().numpy()
    pb, psb = _perturbed(p0, ps0, 4)
    ctx.set(H.HAMR_P, pb)
    ctx.set(H.HAMR_PS, psb)
    ctx.rdump_read_dev(buf.data_ptr())
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    for n in range(ctx.nblk):
        wp, wps = oracle.rdump_read(bs, (1, 1, 1), 1, N2, rec[off[n]:off[n + 1]], g[n], _flat(pb[n]),
                                    _flat(psb[n]))
        assert np.array_equal(_flat(gp[n]), wp) and np.array_equal(_flat(gps[n]), wps), n
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)


def test_file_round_trip(uni, tmp_path):
    """restart_write then restart_read (read + ghost fill) over a perturbed
    state.  The face field is perturbed only where the read or the fill writes
    it: a few ghost faces outside [NG, NG+BS] are written by neither (as in the
    reference) and keep whatever they held."""
    ctx, p0, ps0, g = uni
    ctx.rdump_write(str(tmp_path))
    pb, psb = _perturbed(p0, ps0, 5)
    keep = np.ones(ctx.D, dtype=bool)
    keep[NG:-NG + 1, NG:-NG + 1, NG:-NG + 1] = False
    psb[:, :, keep] = ps0[:, :, keep]
    ctx.set(H.HAMR_P, pb)
    ctx.set(H.HAMR_PS, psb)
    ctx.rdump_read(str(tmp_path))
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    assert np.array_equal(gp, p0)
    np.testing.assert_allclose(gps, ps0, rtol=1e-15, atol=0)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)


def test_bad_arguments_fail(uni):
    ctx, p0, ps0, g = uni
    with pytest.raises(H.HamrError):
        ctx.rdump_block_read(0, np.zeros(ctx.rdump_size(0)), (0, 1, 1))
    with pytest.raises(H.HamrError):
        ctx.rdump_block_read(0, np.zeros(10))
    with pytest.raises(H.HamrError):
        ctx.rdump_block_write(ctx.nblk)


def test_multilevel_records_match_oracle(oracle):
    desc = H.make_desc((64, 32, 32), (1, 2, 1))
    amr = H.make_amr_desc([((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 128))], seed=5)
    ctx = H.AmrContext(desc, amr)
    try:
        ctx.set_grid()
        ctx.init_torus(H.torus_desc())
        ctx.boundprim(1)
        p0 = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_P)]
        ps0 = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_PS)]
        recs = []
        for n in range(ctx.nblk):
            row = ctx.layout[n]
            bs = [int(b) for b in row[4:7]]
            D = ctx.Ds[n]
            gd = _gdet(ctx, int(row[7]), D[0], D[1])
            rec = ctx.rdump_block_write(n)
            assert np.array_equal(rec, oracle.rdump_write(bs, _flat(p0[n]), _flat(ps0[n]), gd)), n
            recs.append((bs, gd, rec))
        off = ctx.rdump_offsets()
        buf = torch.zeros(int(off[-1]), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ctx.rdump_write_dev(buf.data_ptr())
        assert np.array_equal(buf.cpu().numpy(), np.concatenate([r[2] for r in recs]))
        pb = [a * 1.25 for a in p0]
        psb = [a - 0.5 for a in ps0]
        ctx.set(H.HAMR_P, pb)
        ctx.set(H.HAMR_PS, psb)
        ctx.rdump_read_dev(buf.data_ptr())
        gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
        for n, (bs, gd, rec) in enumerate(recs):
            wp, wps = oracle.rdump_read(bs, (1, 1, 1), 1, 2 * 32, rec, gd, _flat(pb[n]), _flat(psb[n]))
            assert np.array_equal(_flat(gp[n]), wp) and np.array_equal(_flat(gps[n]), wps), n
        print("multi-level records of %d blocks bit-exact" % ctx.nblk, flush=True)
    finally:
        ctx.close()


def _centre_metric(ctx, geo, D1, D2):
    m = ctx.download_metric(geo)
    ncol = D1 * D2
    return (np.ascontiguousarray(m["gcov"][H.CENT, :, :ncol].T), np.ascontiguousarray(m["gcon"][H.CENT, :, :ncol].T))


def test_dump_records_match_oracle(uni, oracle):
    """dump_block records (9 floats per interior cell) bit-exact against the
    oracle; the whole-grid device write equals the per-block writes."""
    ctx, p0, ps0, g = uni
    bs = [d - 2 * NG for d in ctx.D]
    recs = []
    for n in range(ctx.nblk):
        gc, gi = _centre_metric(ctx, int(ctx.block_row(n)[0]), ctx.D[0], ctx.D[1])
        rec = ctx.dump_block_write(n)
        assert np.array_equal(rec, oracle.dump_block(bs, _flat(p0[n]), gc, gi)), n
        recs.append(rec)
    off = ctx.dump_offsets()
    buf = torch.zeros(int(off[-1]), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.dump_write_dev(buf.data_ptr())
    assert np.array_equal(buf.cpu().numpy(), np.concatenate(recs))
    print("dump records of %d blocks bit-exact" % ctx.nblk, flush=True)


def test_resume_from_restart_files(uni, tmp_path):
    """restart_write after two steps, a fresh context restart_read from the
    files, two more steps: the state of the uninterrupted four steps to
    rounding (the face field passes through B gdet / gdet)."""
    ctx, p0, ps0, g = uni
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    dt = ctx.estimate_dt()
    dt = ctx.step(2, dt)
    ctx.rdump_write(str(tmp_path))
    dt_mid = dt
    ctx.step(2, dt)
    want_p, want_ps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    c2 = H.Context(desc)
    try:
        c2.set_grid()
        c2.rdump_read(str(tmp_path))
        c2.step(2, dt_mid)
        got_p, got_ps = c2.get(H.HAMR_P), c2.get(H.HAMR_PS)
    finally:
        c2.close()
    I = slice(NG, -NG)
    for k in range(NPR):
        err = np.abs(got_p[:, k, I, I, I] - want_p[:, k, I, I, I]).max()
        scale = np.abs(want_p[:, k, I, I, I]).max()
        assert err <= 1e-9 * scale, (k, err, scale)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
