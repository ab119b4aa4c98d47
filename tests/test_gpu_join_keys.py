"""Join keys beyond one INT32/INT64 lane on the device (MI355X) against the CPU oracle: key tuples
over more than 64 bits (two INT64, INT32 + INT64 + DECIMAL128, four mixed lanes), a DECIMAL128 or
FLOAT64 key alone (compared bit for bit) and UTF8 keys -- plain or dictionary-encoded, in host or
device memory -- mapped to ids by the build side's device dictionary (GroupKeyTable.lookup,
kernels/src/main/java/io/vecruntime/kernels/GroupKeyTable.java:772-826; toIds :1308-1430).

Parity bar: bit-exact pairs in the reference's order (streamed rows in order, each row's build rows
in build order), semi/anti rows and payload outputs; the oracle is pinned against a nested loop in
tests/test_oracle.py."""
import numpy as np
import pytest

from join_key_cases import KIND_SETS, make_case
from oracle.oracle import OracleJoin
from vecruntime_amd import abi
from vecruntime_amd.columns import Column

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def on_device(col, keep):
    """The host Column's buffers copied to device tensors (a dictionary stays on the host)."""
    import torch

    def t(a):
        if a is None:
            return None
        x = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        keep.append(x)
        return x
    return Column(col.type, t(col.data), t(col.validity), offsets=t(col.offsets), dictionary=col.dictionary,
                  length=col.length, mem=abi.MEM_DEVICE)


def seed_of(kinds):
    return 1000 + sum(map(ord, "".join(kinds)))


@pytest.mark.parametrize("kinds", KIND_SETS)
def test_keys_match_oracle(dev, kinds):
    from vecruntime_amd import vr
    case = make_case(seed_of(kinds), kinds, 20_000, 60_000, 4000, dict_probe="utf8" in kinds)
    nk, nb, npr = case["nk"], case["nb"], case["npr"]
    keys = list(range(nk))
    oj = OracleJoin(case["spec"], case["build"], nb)
    bt = vr.BuildTable(dev, case["spec"])
    bt.append(case["build"][:], nb)
    bt.finish()
    assert bt.info()[1] == oj.distinct_keys()
    cap = npr * 40
    for jt in (abi.INNER, abi.OUTER, abi.LEFT_SEMI, abi.LEFT_ANTI):
        pairs = jt in (abi.INNER, abi.OUTER)
        wp, wb, wouts, wr = oj.probe(case["probe"], keys, npr, jt, probe_out=[nk], build_out=[nk] if pairs else [],
                                     capacity=cap)
        gp, gb, gouts, gr = bt.probe(case["probe"], keys, npr, jt, probe_out=[nk], build_out=[nk] if pairs else [],
                                     capacity=cap)
        assert gr == wr and (gr > 0 or jt == abi.LEFT_ANTI)
        assert gp.tolist() == wp[:wr].tolist()
        if pairs:
            assert gb.tolist() == wb[:wr].tolist()
        for g, w in zip(gouts, wouts):
            assert g.values(gr) == w.values(wr)
    wp, _, _, wr = oj.probe(case["probe"], keys, npr, abi.LEFT_ANTI, drop_null_keys=True, capacity=cap)
    gp, _, _, gr = bt.probe(case["probe"], keys, npr, abi.LEFT_ANTI, drop_null_keys=True, capacity=cap)
    assert gp.tolist() == wp[:wr].tolist()
    bt.close()


@pytest.mark.parametrize("kinds", [("utf8",), ("utf8", "int32", "int64", "decimal"), ("int32", "int64", "decimal")])
def test_device_resident_build_and_batched_probe(dev, kinds):
    """The build side appended from device memory in two pieces (plain UTF8 on the device), the
    probe side in device-resident batches through one vr_join_probe_batches launch."""
    import torch
    from vecruntime_amd import vr
    case = make_case(seed_of(kinds) + 1, kinds, 30_000, 90_000, 6000, dict_probe=False)
    nk, nb, npr = case["nk"], case["nb"], case["npr"]
    keep = []
    oj = OracleJoin(case["spec"], case["build"], nb)
    bt = vr.BuildTable(dev, case["spec"])
    half = nb // 2
    for lo, hi in ((0, half), (half, nb)):
        part = [slice_col(c, lo, hi) for c in case["build"]]
        bt.append([on_device(c, keep) for c in part], hi - lo)
    bt.finish()
    cuts = [0, 17, 40_000, npr]
    batches = []
    for lo, hi in zip(cuts, cuts[1:]):
        batches.append(([on_device(slice_col(c, lo, hi), keep) for c in case["probe"]], hi - lo))
    pb = vr.ProbeBatches(batches)
    for jt in (abi.INNER, abi.LEFT_SEMI, abi.LEFT_ANTI):
        wp, wb, _, wr = oj.probe(case["probe"], list(range(nk)), npr, jt, capacity=npr * 40)
        cap = max(1, wr)
        pidx = torch.empty(cap, dtype=torch.int32, device="cuda")
        bidx = torch.empty(cap, dtype=torch.int32, device="cuda") if jt == abi.INNER else None
        rows, per = bt.probe_batches(pb, list(range(nk)), jt, capacity=cap, probe_idx=pidx, build_idx=bidx)
        assert rows == wr
        # batch-relative streamed rows back to the whole probe side's numbering
        base = np.repeat(np.array(cuts[:-1]), per)
        assert (pidx[:rows].cpu().numpy() + base).tolist() == wp[:wr].tolist()
        if jt == abi.INNER:
            assert bidx[:rows].cpu().numpy().tolist() == wb[:wr].tolist()
    bt.close()


def slice_col(col, lo, hi):
    """Rows [lo, hi) of a host Column (bitmaps re-packed, offsets rebased)."""
    from vecruntime_amd.columns import pack_bits, unpack_bits
    n = hi - lo
    valid = None if col.validity is None else pack_bits(unpack_bits(col.validity, col.length)[lo:hi])
    if col.type == abi.UTF8 and col.dictionary is None:
        offs = col.offsets[lo:hi + 1]
        data = col.data[offs[0]:offs[-1]] if offs[-1] > offs[0] else np.zeros(1, np.uint8)
        return Column(abi.UTF8, np.ascontiguousarray(data), valid, offsets=(offs - offs[0]).astype(np.int32), length=n)
    if col.type == abi.DECIMAL128:
        return Column(abi.DECIMAL128, np.ascontiguousarray(col.data[2 * lo:2 * hi]), valid, length=n)
    return Column(col.type, np.ascontiguousarray(col.data[lo:hi]), valid, dictionary=col.dictionary, length=n)
