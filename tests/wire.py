"""Reader of the reference's wire format (src/core/serialization.zig, docs/wire.md)
for Fri.Proof, used to check the bytes field by field: little-endian unsigned
ints (usize as 8 bytes), digests verbatim, slices as a u64 length then the
elements, QM31 as four u32; trailing or missing bytes are errors."""
import struct


class WireError(ValueError):
    pass


class Reader:
    def __init__(self, b):
        self.b, self.pos = b, 0

    def take(self, n):
        if self.pos + n > len(self.b):
            raise WireError("UnexpectedEnd")
        v = self.b[self.pos:self.pos + n]
        self.pos += n
        return v

    def uint(self, n):
        return int.from_bytes(self.take(n), "little")

    def qm31(self):
        return list(struct.unpack("<4I", self.take(16)))

    def digests(self):
        n = self.uint(8)
        return [self.take(32) for _ in range(n)]


def parse_fri_proof(b):
    r = Reader(b)
    out = {"log_size": r.uint(1), "log_blowup": r.uint(1), "num_queries": r.uint(8), "remainder_log": r.uint(1)}
    out["roots"] = r.digests()
    out["remainder"] = [r.qm31() for _ in range(r.uint(8))]
    queries = []
    for _ in range(r.uint(8)):
        q = {"index": r.uint(8), "pairs": []}
        for _ in range(r.uint(8)):
            q["pairs"].append({"value0": r.qm31(), "value1": r.qm31(), "path0": r.digests(), "path1": r.digests()})
        queries.append(q)
    out["queries"] = queries
    if r.pos != len(b):
        raise WireError("TrailingBytes")
    return out
