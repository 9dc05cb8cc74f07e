"""The AMR transfer operators on the device against the oracle, through the
C-ABI: the reference's single-call forms (faceUnknowns*/volumeUnknowns*
Prolongation/Restriction) at orders 1..5 over one and two levels, and the
batched device forms over whole coarse/fine block sets, restriction
accumulating several children in list order.  Bar: 1e-12 of each array's
scale (the device sums the tensor-product operator direction by direction,
the reference node by node: the same linear map, other rounding).
"""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _solvers(exa, oracle, order, nvar=5, npar=1):
    return exa.euler_solver(order, npar=npar), oracle.make_solver(order, nvar=nvar, npar=npar)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("level_delta", [1, 2])
def test_single_volume_transfers_vs_oracle(exa, oracle, order, level_delta):
    print(f"case: volume prolongation/restriction N={order} levels={level_delta}", flush=True)
    rng = np.random.default_rng(order * 7 + level_delta)
    s, os_ = _solvers(exa, oracle, order)
    B, D = order + 1, 6
    n = 3 ** level_delta
    for sub in [(0, 0, 0), (n - 1, 0, 1), tuple(rng.integers(0, n, 3))]:
        uc = rng.standard_normal(B ** 3 * D)
        dev = exa.volume_unknowns_prolongation(s, uc, 1, 1 + level_delta, sub)
        ref = oracle.volume_unknowns_prolongation(os_, uc, 1, 1 + level_delta, sub)
        assert _rel(dev, ref) <= 1e-12, sub
        uf, acc = rng.standard_normal(B ** 3 * D), rng.standard_normal(B ** 3 * D)
        dev = exa.volume_unknowns_restriction(s, acc, uf, 1, 1 + level_delta, sub)
        ref = oracle.volume_unknowns_restriction(os_, acc, uf, 1, 1 + level_delta, sub)
        assert _rel(dev, ref) <= 1e-12, sub


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("level_delta", [1, 2])
def test_single_face_transfers_vs_oracle(exa, oracle, order, level_delta):
    print(f"case: face prolongation/restriction N={order} levels={level_delta}", flush=True)
    rng = np.random.default_rng(100 + order * 7 + level_delta)
    s, os_ = _solvers(exa, oracle, order)
    B, D, V = order + 1, 6, 5
    n = 3 ** level_delta
    for sub in [(0, 0), (n - 1, 1), tuple(rng.integers(0, n, 2))]:
        qc, fc = rng.standard_normal(B * B * D), rng.standard_normal(B * B * V)
        dq, df = exa.face_unknowns_prolongation(s, qc, fc, 0, level_delta, sub)
        rq, rf = oracle.face_unknowns_prolongation(os_, qc, fc, 0, level_delta, sub)
        assert _rel(dq, rq) <= 1e-12 and _rel(df, rf) <= 1e-12, sub
        qf, ff = rng.standard_normal(B * B * D), rng.standard_normal(B * B * V)
        dq, df = exa.face_unknowns_restriction(s, qc, fc, qf, ff, 0, level_delta, sub)
        rq, rf = oracle.face_unknowns_restriction(os_, qc, fc, qf, ff, 0, level_delta, sub)
        assert _rel(dq, rq) <= 1e-12 and _rel(df, rf) <= 1e-12, sub


def test_single_call_level_rules(exa):
    s = exa.euler_solver(3, npar=1)
    uc = np.ones(64 * 6)
    uf = np.full(64 * 6, 7.0)
    # prolongation over zero levels leaves the fine block as it was (the reference's loop does not run)
    assert np.array_equal(exa.volume_unknowns_prolongation(s, uc, 2, 2, (0, 0, 0), uF=uf), uf)
    with pytest.raises(exa.ExaError):  # restriction needs a finer level
        exa.volume_unknowns_restriction(s, uc, uf, 2, 2, (0, 0, 0))
    with pytest.raises(exa.ExaError):  # subinterval index beyond 3^levels
        exa.volume_unknowns_prolongation(s, uc, 0, 1, (3, 0, 0))


def _children(nc, level_delta):
    """Fine blocks of an nc^3 coarse grid refined by 3^level_delta: items
    (fine, coarse, sx, sy, sz) in coarse-major order, and the CSR pointer."""
    r = 3 ** level_delta
    nf = nc * r
    items, ptr = [], [0]
    for c in range(nc ** 3):
        ci, cj, ck = c % nc, (c // nc) % nc, c // (nc * nc)
        for sz, sy, sx in itertools.product(range(r), repeat=3):
            fi, fj, fk = ci * r + sx, cj * r + sy, ck * r + sz
            items.append((fi + nf * (fj + nf * fk), c, sx, sy, sz))
        ptr.append(len(items))
    return np.array(items, dtype=np.int32), np.array(ptr, dtype=np.int32)


@pytest.mark.parametrize("order,level_delta", [(3, 1), (2, 2), (5, 1)])
def test_volume_batches_vs_oracle(exa, oracle, order, level_delta):
    print(f"case: volume batches N={order} levels={level_delta}", flush=True)
    rng = np.random.default_rng(order + 10 * level_delta)
    s, os_ = _solvers(exa, oracle, order)
    B, D = order + 1, 6
    nc = 2
    items, ptr = _children(nc, level_delta)
    nfine = len(items)
    coarse = rng.standard_normal((nc ** 3, B ** 3 * D))
    d_coarse, d_fine = exa.DeviceArray(coarse), exa.DeviceArray(np.zeros((nfine, B ** 3 * D)))
    # prolongation items: (fine destination, coarse source, sub); restriction: (coarse, fine, sub)
    d_items, d_ptr = exa.DeviceArray(items), exa.DeviceArray(ptr)
    d_ritems = exa.DeviceArray(items[:, [1, 0, 2, 3, 4]])
    ms = exa.volume_prolongation_batch(s, d_fine, d_coarse, d_items, level_delta)
    assert ms > 0.0
    fine = d_fine.get()
    for it in items[:: max(1, nfine // 40)]:
        ref = oracle.volume_unknowns_prolongation(os_, coarse[it[1]], 0, level_delta, it[2:])
        assert _rel(fine[it[0]], ref) <= 1e-12
    # restriction: every coarse block gets its children added in list order
    fine_in = rng.standard_normal((nfine, B ** 3 * D))
    start = rng.standard_normal((nc ** 3, B ** 3 * D))
    d_fine_in, d_acc = exa.DeviceArray(fine_in), exa.DeviceArray(start)
    exa.volume_restriction_batch(s, d_acc, d_fine_in, d_ptr, d_ritems, level_delta)
    acc = d_acc.get()
    for c in range(nc ** 3):
        ref = start[c].copy()
        for it in items[ptr[c]:ptr[c + 1]]:
            ref = oracle.volume_unknowns_restriction(os_, ref, fine_in[it[0]], 0, level_delta, it[2:])
        assert _rel(acc[c], ref) <= 1e-12, c


def test_face_batches_vs_oracle(exa, oracle):
    print("case: face batches N=4 one level", flush=True)
    rng = np.random.default_rng(44)
    s, os_ = _solvers(exa, oracle, 4)
    B, D, V = 5, 6, 5
    ncf = 4  # coarse face blocks, each split into 9 fine face blocks
    items = np.array([(9 * c + 3 * b + a, c, a, b, 0) for c in range(ncf) for b in range(3) for a in range(3)],
                     dtype=np.int32)
    ptr = np.arange(0, 9 * ncf + 1, 9, dtype=np.int32)
    qc, fc = rng.standard_normal((ncf, B * B * D)), rng.standard_normal((ncf, B * B * V))
    d_qc, d_fc = exa.DeviceArray(qc), exa.DeviceArray(fc)
    d_qf, d_ff = exa.DeviceArray(np.zeros((9 * ncf, B * B * D))), exa.DeviceArray(np.zeros((9 * ncf, B * B * V)))
    d_items, d_ptr = exa.DeviceArray(items), exa.DeviceArray(ptr)
    d_ritems = exa.DeviceArray(items[:, [1, 0, 2, 3, 4]])
    exa.face_prolongation_batch(s, d_qf, d_ff, d_qc, d_fc, d_items, 1)
    qf, ff = d_qf.get(), d_ff.get()
    for it in items:
        rq, rf = oracle.face_unknowns_prolongation(os_, qc[it[1]], fc[it[1]], 0, 1, it[2:4])
        assert _rel(qf[it[0]], rq) <= 1e-12 and _rel(ff[it[0]], rf) <= 1e-12
    exa.face_restriction_batch(s, d_qc, d_fc, d_qf, d_ff, d_ptr, d_ritems, 1)
    q2, f2 = d_qc.get(), d_fc.get()
    for c in range(ncf):
        rq, rf = qc[c].copy(), fc[c].copy()
        for it in items[ptr[c]:ptr[c + 1]]:
            rq, rf = oracle.face_unknowns_restriction(os_, rq, rf, qf[it[0]], ff[it[0]], 0, 1, it[2:4])
        assert _rel(q2[c], rq) <= 1e-12 and _rel(f2[c], rf) <= 1e-12
    # prolongation then complete restriction onto zero blocks returns the coarse data
    z_q, z_f = exa.DeviceArray(np.zeros_like(qc)), exa.DeviceArray(np.zeros_like(fc))
    exa.face_restriction_batch(s, z_q, z_f, d_qf, d_ff, d_ptr, d_ritems, 1)
    assert _rel(z_q.get(), qc) <= 1e-11 and _rel(z_f.get(), fc) <= 1e-11


def test_batch_rejects_out_of_range_items(exa):
    s = exa.euler_solver(2, npar=1)
    blk = 27 * 6
    d_c, d_f = exa.DeviceArray(np.ones((2, blk))), exa.DeviceArray(np.zeros((3, blk)))
    bad = exa.DeviceArray(np.array([(0, 0, 0, 0, 0), (5, 0, 0, 0, 0), (1, 1, 3, 0, 0)], dtype=np.int32))
    with pytest.raises(exa.ExaError):
        exa.volume_prolongation_batch(s, d_f, d_c, bad, 1)
    assert np.allclose(d_f.get()[0], 1.0)  # the valid item ran
    bad_ptr = exa.DeviceArray(np.array([0, 7], dtype=np.int32))  # a range beyond the item list
    good = exa.DeviceArray(np.array([(0, 0, 0, 0, 0)], dtype=np.int32))
    with pytest.raises(exa.ExaError):
        exa.volume_restriction_batch(s, d_c, d_f, bad_ptr, good, 1)


def test_batch_rejects_repeated_destinations(exa, oracle):
    """A destination block named twice in one call (two prolongation items, or
    two restriction ranges) is written by one transfer only and the call fails."""
    print("case: repeated AMR destinations", flush=True)
    s, os_ = _solvers(exa, oracle, 2)
    blk = 27 * 6
    rng = np.random.default_rng(5)
    coarse = rng.standard_normal((2, blk))
    d_c, d_f = exa.DeviceArray(coarse), exa.DeviceArray(np.zeros((3, blk)))
    dup = exa.DeviceArray(np.array([(1, 0, 0, 0, 0), (2, 1, 1, 1, 1), (1, 1, 2, 0, 1)], dtype=np.int32))
    with pytest.raises(exa.ExaError):
        exa.volume_prolongation_batch(s, d_f, d_c, dup, 1)
    fine = d_f.get()
    assert np.array_equal(fine[0], np.zeros(blk))  # named by no item
    ref2 = oracle.volume_unknowns_prolongation(os_, coarse[1], 0, 1, (1, 1, 1))
    assert _rel(fine[2], ref2) <= 1e-12  # the unique destination is written
    cands = [oracle.volume_unknowns_prolongation(os_, coarse[0], 0, 1, (0, 0, 0)),
             oracle.volume_unknowns_prolongation(os_, coarse[1], 0, 1, (2, 0, 1))]
    assert min(_rel(fine[1], c) for c in cands) <= 1e-12  # exactly one of the two transfers landed
    # restriction: two ranges with the same destination
    start = rng.standard_normal((2, blk))
    d_acc = exa.DeviceArray(start)
    ritems = exa.DeviceArray(np.array([(0, 0, 0, 0, 0), (0, 1, 1, 0, 0), (1, 2, 2, 2, 2)], dtype=np.int32))
    rptr = exa.DeviceArray(np.array([0, 1, 2, 3], dtype=np.int32))
    d_fin = exa.DeviceArray(rng.standard_normal((3, blk)))
    with pytest.raises(exa.ExaError):
        exa.volume_restriction_batch(s, d_acc, d_fin, rptr, ritems, 1)
    # the same ranges with distinct destinations are accepted
    ok_items = exa.DeviceArray(np.array([(0, 0, 0, 0, 0), (0, 1, 1, 0, 0), (1, 2, 2, 2, 2)], dtype=np.int32))
    ok_ptr = exa.DeviceArray(np.array([0, 2, 3], dtype=np.int32))
    exa.volume_restriction_batch(s, exa.DeviceArray(start), d_fin, ok_ptr, ok_items, 1)
