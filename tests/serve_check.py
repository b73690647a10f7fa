"""Host restatement of the serving queries' results (count and checksum,
the terms of include/pimdal_gen.h), for the tests.  The select checksum
depends on the output order (row id at each output position); the join
and group-by checksums do not."""
import numpy as np

M64 = (1 << 64) - 1
CK_SELECT, CK_JOIN, CK_GROUP = 0x53454C4543540001, 0x4A4F494E00000002, 0x47524F5550000003


def mix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x ^ (x >> np.uint64(30))
        x = x * np.uint64(0xBF58476D1CE4E5B9)
        x = x ^ (x >> np.uint64(27))
        x = x * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def mix64_int(x: int) -> int:
    return int(mix64(np.array([x & M64], dtype=np.uint64))[0])


def usum(a) -> int:
    return int(np.sum(np.asarray(a, dtype=np.uint64), dtype=np.uint64)) if len(a) else 0


def ck_select_ordered(rows) -> int:
    """Checksum of selected fact rows in output order (row at position i)."""
    rows = np.asarray(rows, dtype=np.uint64)
    pos = np.arange(rows.size, dtype=np.uint64)
    return usum(mix64(((pos << np.uint64(32)) | rows) ^ np.uint64(CK_SELECT)))


def ck_join(fact_rows, dim_rows) -> int:
    fr = np.asarray(fact_rows, dtype=np.uint64)
    dr = np.asarray(dim_rows, dtype=np.uint64)
    return usum(mix64(((fr << np.uint64(32)) | dr) ^ np.uint64(CK_JOIN)))


def expected(q, fact, dim, orc, dim_sorted=None):
    kind, off, rows, param = int(q["kind"]), int(q["offset"]), int(q["rows"]), int(q["param"])
    sl = slice(off, off + rows)
    if kind == 1:
        r = off + np.nonzero(fact["price"][sl] < param)[0].astype(np.uint64)
        return len(r), ck_select_ordered(r)
    if kind == 2:
        order, skeys = dim_sorted
        k = fact["orderkey"][sl]
        pos = np.searchsorted(skeys, k)
        pos = np.minimum(pos, len(skeys) - 1)
        found = skeys[pos] == k
        drow = order[pos]
        ok = found & (dim["date"][drow] < param)
        fr = (off + np.nonzero(ok)[0]).astype(np.uint64)
        dr = drow[ok].astype(np.uint64)
        return len(fr), usum(mix64(((fr << np.uint64(32)) | dr) ^ np.uint64(CK_JOIN)))
    t = {c: fact[c][sl] for c in ("date", "flag1", "flag2", "quantity", "price", "discount", "tax")}
    g = orc.q1_flags(t, param)
    ck = 0
    for i in range(len(g["key"])):
        h = mix64_int(int(g["key"][i]) ^ CK_GROUP)
        ch = g["sum_charge"][i]
        for v in (g["sum_qty"][i], g["sum_price"][i], g["sum_disc"][i], g["sum_disc_price"][i],
                  ch & M64, (ch >> 64) & M64, g["count"][i]):
            h = mix64_int(h ^ (int(v) & M64))
        ck = (ck + h) & M64
    return len(g["key"]), ck


def expected_oracle(q, fact, dim, orc):
    """The same (count, checksum) through the CPU restatement: the oracle's
    order-preserving select of the fact rows (row ids in output order), the
    oracle's hash join of the qualifying dimension rows with the fact slice
    (pairs in any order), and its Q1 group-by of the slice."""
    kind, off, rows, param = int(q["kind"]), int(q["offset"]), int(q["rows"]), int(q["param"])
    sl = slice(off, off + rows)
    if kind == 1:
        _, rid = orc.select(fact["price"][sl], np.zeros(rows, np.uint32), -(1 << 63), param)
        return len(rid), ck_select_ordered(rid.astype(np.uint64) + np.uint64(off))
    if kind == 2:
        dpay, drow = orc.select(dim["date"], dim["key"].view(np.uint32), -(1 << 31), param)
        fk = fact["orderkey"][sl].view(np.uint32)
        frows = np.arange(off, off + rows, dtype=np.uint32)
        fr, dr = orc.hash_join(dpay, drow, fk, frows)
        return len(fr), ck_join(fr, dr)
    return expected(q, fact, dim, orc)
