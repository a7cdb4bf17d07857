"""Decode the reference's solver fixture into a compact numpy archive.

Source: ``algo/linear_solver/test/lin_prob.xml.bz2`` of SternheimerGW (read by
``linear_problem_read`` in ``algo/linear_solver/test/linear_solver.pf:44-97``).
The file is an iotk binary XML stream.  Each data tag ``<NAME ... size=N
kind=K>`` is followed by ``\\n``, the 4-byte trailer of the tag record, a
4-byte record length L, a 4-byte dummy, and L-4 bytes of little-endian
payload.  Complex data are (re, im) float64 pairs, matrices column-major.

The output ``lin_prob.npz`` holds data only: DIMENSION, NUMBER_SHIFT,
LIST_SHIFT (70), LINEAR_OPERATOR (283x283, column-major order restored),
RIGHT_HAND_SIDE (283), INCORRECT_SOLUTION (283x70).  Run in a container
where the reference tree is mounted; the archive is committed.
"""
import bz2
import hashlib
import os
import sys

import numpy as np

SRC = "/root/reference/algo/linear_solver/test/lin_prob.xml.bz2"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lin_prob.npz")


def _record(data, tag):
    start = data.find(b"<" + tag.encode() + b" ")
    if start < 0:
        raise KeyError(tag)
    close = data.find(b">", start)
    head = data[start:close + 1].decode()
    attrs = dict(kv.split("=") for kv in head[1:-1].split()[1:])
    attrs = {k: v.strip('"') for k, v in attrs.items()}
    pos = close + 1
    assert data[pos:pos + 1] == b"\n"
    pos += 1 + 4
    reclen = int.from_bytes(data[pos:pos + 4], "little")
    pos += 4 + 4
    payload = data[pos:pos + reclen - 4]
    size = int(attrs["size"])
    kind = int(attrs["kind"])
    if attrs["type"] == "integer":
        arr = np.frombuffer(payload, dtype="<i%d" % kind, count=size)
    elif attrs["type"] == "complex":
        arr = np.frombuffer(payload, dtype="<c%d" % (2 * kind), count=size)
    else:
        raise ValueError(attrs["type"])
    return arr.copy()


def decode(path=SRC):
    data = bz2.open(path, "rb").read()
    n = int(_record(data, "DIMENSION")[0])
    ns = int(_record(data, "NUMBER_SHIFT")[0])
    sigma = _record(data, "LIST_SHIFT")
    amat = _record(data, "LINEAR_OPERATOR").reshape((n, n), order="F")
    bb = _record(data, "RIGHT_HAND_SIDE")
    xx = _record(data, "INCORRECT_SOLUTION").reshape((n, ns), order="F")
    return dict(DIMENSION=n, NUMBER_SHIFT=ns, LIST_SHIFT=sigma,
                LINEAR_OPERATOR=amat, RIGHT_HAND_SIDE=bb,
                INCORRECT_SOLUTION=xx,
                source_sha256=hashlib.sha256(open(path, "rb").read()).hexdigest())


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    d = decode(src)
    np.savez_compressed(OUT, **{k: np.asarray(v) for k, v in d.items()})
    print("wrote", OUT, os.path.getsize(OUT), "bytes; n =", d["DIMENSION"],
          "shifts =", d["NUMBER_SHIFT"])
    # digest of the decompressed stream: pins a writer of the format byte for byte
    raw = bz2.open(src, "rb").read()
    stream = os.path.join(os.path.dirname(OUT), "lin_prob_stream.json")
    with open(stream, "w") as f:
        f.write('{"bytes": %d, "sha256": "%s"}\n' % (len(raw), hashlib.sha256(raw).hexdigest()))
    print("wrote", stream)
