"""The symmetric mixed-precision pair kernel (each SI x SI pair evaluated once
and applied to both bodies, the reference CPU path's i < j scheme,
pp_disk.cu:822-865) against the FP64 kernel, which is bit-identical to the
oracle (test_gpu_parity.py): error bounds at every chunking the engine uses,
independence of the launch shape, and agreement with the one-sided kernel."""
import numpy as np
import pytest

import redamd as R

pytestmark = pytest.mark.gpu


def acc(d, r, v):
    a = d.calc_dydx(1, 0, 0.0, r, v)
    return np.stack([a["x"], a["y"], a["z"]], 1)


def fp64_and_mutual(r, v, p, b, nbt):
    """FP64 accelerations, and the mutual part alone (star mass 0)."""
    d = R.Disk(r, v, p, b, nbt)
    ref = acc(d, r, v)
    p0 = p.copy()
    p0["mass"][0] = 0.0
    mut = acc(R.Disk(r, v, p0, b, nbt), r, v)
    return ref, mut


def check_bounds(got, ref, mut):
    """FP32 pair terms (2^-22 each, rsqrt included) in FP32 partials of at most
    256 terms folded in FP64: the error is bounded relative to the mutual
    acceleration (the star's pull on the disk is FP64 and exact; the star
    itself feels the mutual sum alone, mut[0] = ref[0])."""
    err = np.linalg.norm(got - ref, axis=1)
    m = np.linalg.norm(mut, axis=1)
    assert np.all(err <= 2e-5 * m + 1e-18), (err / m).max()


# bodies -> 256-body blocks -> chunks of R blocks: 201 -> one block whose
# fourth row is partial, 256 -> one full block (the diagonal tiles of disks of
# at most 24 blocks take each pair once with its reaction), 151 -> one block of
# three rows (one-sided diagonal tile), 301 -> 2 blocks (partial),
# 5000 -> 20 (R = 1), 70000 -> 274 (R = 2, 137 chunks), 262200 -> 1025 (R = 8),
# 1048576 -> 4096 (R at its cap of 8: 512 chunks, past the 255 the cap-free
# rule allows, so 512 partial-sum slots per body)
@pytest.mark.parametrize("n,seed", [(200, 37), (255, 38), (150, 39), (300, 31), (4999, 32), (69999, 33),
                                    (262199, 34), (1048575, 36)])
def test_symmetric_force_bounds(n, seed):
    r, v, p, b, nbt = R.make_disk(n, seed)
    ref, mut = fp64_and_mutual(r, v, p, b, nbt)
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    d.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC)
    check_bounds(acc(d, r, v), ref, mut)


def test_symmetric_independent_of_launch_shape():
    r, v, p, b, nbt = R.make_disk(69999, 35)
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    outs = []
    for bpc in (1, 3, 16):
        d.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC, blocks_per_cu=bpc)
        outs.append(acc(d, r, v))
        outs.append(acc(d, r, v))
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))


def test_symmetric_agrees_with_one_sided():
    r, v, p, b, nbt = R.make_disk(6000, 36)
    ref, mut = fp64_and_mutual(r, v, p, b, nbt)
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    d.set_tuning(pair_mode=R.RD_PAIRS_ONE_SIDED)
    one = acc(d, r, v)
    d.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC)
    sym = acc(d, r, v)
    check_bounds(one, ref, mut)
    check_bounds(sym, ref, mut)
    # the two FP32 paths differ only by rounding: far inside either bound
    dm = np.linalg.norm(sym - one, axis=1)[1:]
    assert np.all(dm <= 4e-5 * np.linalg.norm(mut[1:], axis=1) + 1e-18)


def test_symmetric_batch_matches_single_disks():
    # uneven disks, each with its own chunking (R = 1 and R = 2), in one launch;
    # disks of at most 24 blocks (6144 bodies) take their diagonal tiles with
    # reactions, so the batch runs the kernel instance with that sweep while
    # the 70000- and 6145-body disks alone run the one without it
    sizes = [(69999, 41), (700, 42), (5000, 43), (129, 44), (6143, 45), (6144, 46)]
    parts = [R.make_disk(n, s) for n, s in sizes]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    bd = R.Disk(*cat, np.stack([q[4] for q in parts]), batch=True)
    bd.set_precision(R.RD_PREC_MIXED)
    bd.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC)
    for k, (r, v, p, b, nbt) in enumerate(parts):
        d = R.Disk(r, v, p, b, nbt)
        d.set_precision(R.RD_PREC_MIXED)
        d.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC)
        single = acc(d, r, v)
        a = bd.calc_dydx(1, 0, 0.0, r, v, k=k)
        batch = np.stack([a["x"], a["y"], a["z"]], 1)
        assert np.array_equal(batch.view(np.uint64), single.view(np.uint64)), k


def test_chunk_min_bounds_and_batch_identity():
    # sym_chunk_min = 2: small disks in chunks of two 256-body blocks (fewer
    # partial-sum slots, items of up to 4 tiles, reactions summed in LDS).
    # Within the FP32 bound of the FP64 kernel, and a batch with that tuning
    # is bit-identical to standalone disks with that tuning.
    sizes = [(700, 51), (5000, 52), (129, 53), (2000, 54)]
    parts = [R.make_disk(n, s) for n, s in sizes]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    bd = R.Disk(*cat, np.stack([q[4] for q in parts]), batch=True)
    bd.set_precision(R.RD_PREC_MIXED)
    bd.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC, sym_chunk_min=2)
    assert bd.tuning().sym_chunk_min == 2
    for k, (r, v, p, b, nbt) in enumerate(parts):
        ref, mut = fp64_and_mutual(r, v, p, b, nbt)
        d = R.Disk(r, v, p, b, nbt)
        d.set_precision(R.RD_PREC_MIXED)
        d.set_tuning(pair_mode=R.RD_PAIRS_SYMMETRIC, sym_chunk_min=2)
        single = acc(d, r, v)
        check_bounds(single, ref, mut)
        a = bd.calc_dydx(1, 0, 0.0, r, v, k=k)
        batch = np.stack([a["x"], a["y"], a["z"]], 1)
        assert np.array_equal(batch.view(np.uint64), single.view(np.uint64)), k


def test_chunk_min_rejects_out_of_range():
    r, v, p, b, nbt = R.make_disk(300, 55)
    d = R.Disk(r, v, p, b, nbt)
    with pytest.raises(R.RedError):
        d.set_tuning(sym_chunk_min=9)
    d.set_tuning(sym_chunk_min=0)          # 0 reads as 1
    assert d.tuning().sym_chunk_min == 1
