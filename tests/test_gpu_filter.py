"""Selection producer and compaction on the device against the oracle (MI355X): vr_filter_compare
(CompareKernels.compareScalar + VectorExpr.selection) and vr_compact (CompactKernels) over
device-resident columns.  Bit-exact: selection bitmaps (padding included), counts, compacted
values, validity, BOOL bits, dictionary codes, UTF8 offsets and bytes."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle
from test_oracle_filter import LITERALS, OPS, column, utf8_column, values
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

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
    """A host Column's buffers copied to the device (the dictionary of an encoded column stays on
    the host)."""
    import torch

    def t(a):
        if a is None:
            return None
        x = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        keep.append(x)
        return x
    return Column(col.type, t(col.data), t(col.validity), offsets=t(col.offsets), dictionary=col.dictionary,
                  length=col.length, mem=abi.MEM_DEVICE)


class DevOut:
    """A device output column: padded validity, data, offsets for UTF8."""

    def __init__(self, t, cap, nbytes=0):
        import torch
        self.type, self.cap = t, cap
        pad = ((cap + 63) // 64 + 1) * 8
        self.valid = torch.zeros(pad, dtype=torch.uint8, device="cuda")
        if t == abi.BOOL:
            self.data = torch.zeros(pad, dtype=torch.uint8, device="cuda")
        elif t == abi.UTF8:
            self.data = torch.zeros(max(1, nbytes), dtype=torch.uint8, device="cuda")
        else:
            w = {abi.INT32: 4, abi.INT64: 8, abi.FLOAT64: 8, abi.DECIMAL128: 16}[t]
            self.data = torch.zeros(max(1, cap) * w, dtype=torch.uint8, device="cuda")
        self.offsets = torch.zeros(cap + 1, dtype=torch.int32, device="cuda") if t == abi.UTF8 else None
        self._c = abi.VrOutColumn(t, abi.MEM_DEVICE, cap, self.valid.data_ptr(), self.data.data_ptr(),
                                  self.offsets.data_ptr() if self.offsets is not None else None)

    def c(self):
        return self._c


def as_bytes(oc, t, m):
    """Comparable host bytes of m compacted rows of an OutColumn (host) or DevOut (device)."""
    if isinstance(oc, DevOut):
        valid = np.unpackbits(oc.valid.cpu().numpy(), bitorder="little")[:m]
        raw = oc.data.cpu().numpy()
        offs = oc.offsets.cpu().numpy()[: m + 1] if oc.offsets is not None else None
    else:
        valid = np.unpackbits(oc.validity, bitorder="little")[:m]
        raw = oc.data.view(np.uint8)
        offs = oc.offsets[: m + 1] if t == abi.UTF8 else None
    if t == abi.BOOL:
        data = np.unpackbits(raw, bitorder="little")[:m].tobytes()
    elif t == abi.UTF8:
        data = raw[: offs[m]].tobytes() + offs.tobytes()
    else:
        w = {abi.INT32: 4, abi.INT64: 8, abi.FLOAT64: 8, abi.DECIMAL128: 16}[t]
        data = raw[: m * w].tobytes()
    return valid.tobytes(), data


@pytest.mark.parametrize("t", [abi.INT32, abi.INT64, abi.FLOAT64, abi.DECIMAL128])
def test_compare_matches_oracle(dev, t):
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(60 + t)
    for n in (1, 63, 64, 65, 1000, 300_007):
        print("compare type %d n %d" % (t, n), flush=True)
        vals = values(rng, t, n)
        nulls = rng.random(n) < 0.15
        sel = pack_bits(rng.random(n) < 0.7)
        host = column(t, vals, nulls)
        keep = []
        dcol = on_device(host, keep)
        dsel = torch.from_numpy(sel).cuda()
        for s in LITERALS[t]:
            for op in OPS:
                for hs, ds in ((None, None), (sel, dsel)):
                    want, wc = oracle.filter_compare(host, op, s, n, hs)
                    got, gc = vr.filter_compare(dev, dcol, op, s, n, ds)
                    assert gc == wc, (t, n, s, op)
                    assert np.array_equal(got.cpu().numpy(), want[: got.numel()]), (t, n, s, op)


@pytest.mark.parametrize("density", [0.0, 0.02, 0.4, 1.0])
def test_compact_matches_oracle(dev, density):
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(int(density * 100) + 7)
    for n in (1, 64, 65, 100_003):
        print("compact density %.2f n %d" % (density, n), flush=True)
        selb = rng.random(n) < density
        sel = pack_bits(selb)
        nulls = [rng.random(n) < 0.2 for _ in range(7)]
        strs = ["".join(chr(97 + int(c)) for c in rng.integers(0, 26, int(k))) for k in rng.integers(0, 9, n)]
        cols = [Column.ints(rng.integers(-100, 100, n), nulls=nulls[0]),
                Column.longs(rng.integers(-10 ** 12, 10 ** 12, n), nulls=nulls[1]),
                Column.doubles(rng.standard_normal(n), nulls=nulls[2]),
                Column.decimals([int(x) * 10 ** 20 + 7 for x in rng.integers(-10 ** 6, 10 ** 6, n)], nulls=nulls[3]),
                Column.bools(rng.random(n) < 0.5, nulls=nulls[4]), utf8_column(strs, nulls[5]),
                utf8_column(["x", "yy", "zzz"], nulls[6], rng.integers(0, 3, n))]
        types = [abi.INT32, abi.INT64, abi.FLOAT64, abi.DECIMAL128, abi.BOOL, abi.UTF8, abi.INT32]
        keep = []
        dcols = [on_device(c, keep) for c in cols]
        dsel = torch.from_numpy(sel).cuda()
        nbytes = vr.compact_utf8_bytes(dev, dcols[5], n, dsel)
        assert nbytes == sum(len(strs[i]) for i in range(n) if selb[i])
        want, m = oracle.compact(cols, n, sel, types, n, utf8_bytes=nbytes)
        outs = [DevOut(t, max(m, 1), nbytes) for t in types]
        got_m = vr.compact(dev, dcols, n, dsel, outs)
        assert got_m == m
        for c, t in enumerate(types):
            assert as_bytes(outs[c], t, m) == as_bytes(want[c], t, m), (c, n, density)


def test_filter_then_compact_and_capacity(dev):
    """The filter operator's sparse branch on one 3M-row batch: int32 < threshold (about 30% pass),
    then three columns compacted, against numpy; a too-small output reports the rows it needs."""
    import torch
    from vecruntime_amd import vr
    n = 3_000_000
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    f = torch.randint(0, 1000, (n,), dtype=torch.int32, device="cuda", generator=g)
    a = torch.randint(-10 ** 9, 10 ** 9, (n,), dtype=torch.int64, device="cuda", generator=g)
    b = torch.rand(n, dtype=torch.float64, device="cuda", generator=g)
    sel, cnt = vr.filter_compare(dev, vr.device_column(abi.INT32, f), abi.CMP_LT, 300, n)
    fh = f.cpu().numpy()
    assert cnt == int((fh < 300).sum())
    outs = [DevOut(abi.INT64, cnt), DevOut(abi.FLOAT64, cnt), DevOut(abi.INT32, cnt)]
    m = vr.compact(dev, [vr.device_column(abi.INT64, a), vr.device_column(abi.FLOAT64, b), vr.device_column(abi.INT32, f)],
                   n, sel, outs)
    assert m == cnt
    keepm = fh < 300
    assert np.array_equal(outs[0].data.cpu().numpy().view(np.int64)[:m], a.cpu().numpy()[keepm])
    assert np.array_equal(outs[1].data.cpu().numpy().view(np.float64)[:m], b.cpu().numpy()[keepm])
    assert np.array_equal(outs[2].data.cpu().numpy().view(np.int32)[:m], fh[keepm])
    small = [DevOut(abi.INT64, 10)]
    with pytest.raises(vr.VrError) as e:
        vr.compact(dev, [vr.device_column(abi.INT64, a)], n, sel, small)
    assert e.value.status == abi.VR_ERR_CAPACITY


def test_gather_matches_oracle(dev):
    """vr_gather of every lane by an index array with -1 padding (a join's probeIdx / buildIdx)."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(33)
    for n, m in ((1, 1), (100, 65), (5000, 200_003)):
        print("gather n %d m %d" % (n, m), flush=True)
        nulls = rng.random(n) < 0.2
        idx = rng.integers(-1, n, m).astype(np.int32)
        strs = ["".join(chr(97 + int(c)) for c in rng.integers(0, 26, int(k))) for k in rng.integers(0, 9, n)]
        cols = [Column.ints(rng.integers(-9, 9, n), nulls=nulls), Column.longs(rng.integers(-10 ** 15, 10 ** 15, n), nulls=nulls),
                Column.doubles(rng.standard_normal(n), nulls=nulls),
                Column.decimals([int(x) * 10 ** 25 - 3 for x in rng.integers(-10 ** 6, 10 ** 6, n)], nulls=nulls),
                Column.bools(rng.random(n) < 0.5, nulls=nulls), utf8_column(strs, nulls)]
        keep = []
        didx = torch.from_numpy(idx).cuda()
        for col in cols:
            dcol = on_device(col, keep)
            nbytes = vr.gather_utf8_bytes(dev, dcol, didx, m) if col.type == abi.UTF8 else 0
            want = oracle.gather(col, idx, utf8_bytes=nbytes)
            out = DevOut(col.type, m, nbytes)
            vr.gather(dev, dcol, didx, m, out)
            assert as_bytes(out, col.type, m) == as_bytes(want, col.type, m), (col.type, n, m)
