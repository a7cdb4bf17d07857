"""Binary iotk LINEAR_PROBLEM files: failure capture and replay of a solve.

When a Green's-function solve misses its residual test, the reference writes
the dense operator, the shifts, the right-hand side and the failed solution
with iotk (phys/green/src/green.f90:437-446), and its solver unit test reads
such a file back (algo/linear_solver/test/linear_solver.pf:44-97).  This
module writes and reads the same binary stream, so a problem captured here
replays in the reference's test and the other way round.

Stream layout (Fortran sequential unformatted records, each framed by its
32-bit length before and after):

* a markup item is two records: a 4-byte control word ``kind | len << 8``
  (kind 5 = ``<?...?>`` instruction, 1 = begin tag, 2 = end tag) and a
  record holding the word ``0x80 | len << 8`` followed by the ``len`` bytes
  of text (``\\n`` + markup + ``\\n``);
* a data item is a begin tag with ``type``, ``size`` and ``kind``
  attributes, one record of a 4-byte zero and the little-endian payload
  (integers; complex as (re, im) float64 pairs, matrices column-major), and
  the end tag.
"""
from __future__ import annotations

import bz2
import struct

import numpy as np

_PI, _BEGIN, _END = 5, 1, 2
_HEADER = ('<?iotk version="1.2.0"?>', '<?iotk file_version="1.0"?>', '<?iotk binary="T"?>',
           '<?iotk qe_syntax="F"?>')
ROOT = "LINEAR_PROBLEM"
FIELDS = ("DIMENSION", "NUMBER_SHIFT", "LIST_SHIFT", "LINEAR_OPERATOR", "RIGHT_HAND_SIDE",
          "INCORRECT_SOLUTION")


def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def _markup(kind: int, text: str) -> bytes:
    body = ("\n" + text + "\n").encode()
    n = len(body)
    return _record(struct.pack("<i", kind | (n << 8))) + _record(
        struct.pack("<i", 0x80 | (n << 8)) + body)


def _data(name: str, arr: np.ndarray, indent: str = "  ") -> bytes:
    arr = np.asarray(arr)
    if np.iscomplexobj(arr):
        typ, kind, payload = "complex", 8, np.asarray(arr, dtype="<c16").tobytes(order="F")
    else:
        typ, kind, payload = "integer", 4, np.asarray(arr, dtype="<i4").tobytes(order="F")
    head = f'{indent}<{name} type="{typ}" size="{arr.size}" kind="{kind}">'
    return (_markup(_BEGIN, head) + _record(b"\x00\x00\x00\x00" + payload)
            + _markup(_END, f"{indent}</{name}>"))


def encode_linear_problem(a, b, sigma, x) -> bytes:
    """The iotk byte stream of one linear problem: a (n, n), b (n,),
    sigma (nshift,), x (n, nshift) complex."""
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[0]
    sigma = np.atleast_1d(np.asarray(sigma, dtype=np.complex128))
    x = np.asarray(x, dtype=np.complex128).reshape(n, sigma.size)
    out = b"".join(_markup(_PI, h) for h in _HEADER)
    out += _markup(_BEGIN, f"<{ROOT}>")
    out += _data("DIMENSION", np.array(n))
    out += _data("NUMBER_SHIFT", np.array(sigma.size))
    out += _data("LIST_SHIFT", sigma)
    out += _data("LINEAR_OPERATOR", a)
    out += _data("RIGHT_HAND_SIDE", np.asarray(b, dtype=np.complex128))
    out += _data("INCORRECT_SOLUTION", x)
    out += _markup(_END, f"</{ROOT}>")
    return out


def write_linear_problem(path: str, a, b, sigma, x) -> None:
    """Write the problem to `path` (bz2-compressed when it ends in .bz2)."""
    data = encode_linear_problem(a, b, sigma, x)
    opener = bz2.open if path.endswith(".bz2") else open
    with opener(path, "wb") as f:
        f.write(data)


def _records(data: bytes):
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        yield data[pos + 4:pos + 4 + n]
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        if m != n:
            raise ValueError(f"iotk: record framing mismatch at byte {pos}")
        pos += 8 + n


def decode_linear_problem(data: bytes) -> dict:
    """Parse a LINEAR_PROBLEM stream into numpy arrays (matrices column-major
    restored)."""
    recs = list(_records(data))
    out, i = {}, 0
    while i < len(recs):
        (ctrl,) = struct.unpack_from("<i", recs[i], 0)
        kind = ctrl & 0xFF
        text = recs[i + 1][4:].decode().strip()
        i += 2
        if kind == _BEGIN and "type=" in text:
            name = text[1:].split()[0]
            attrs = dict(kv.split("=") for kv in text[1:-1].split()[1:])
            attrs = {k: v.strip('"') for k, v in attrs.items()}
            payload = recs[i][4:]
            size = int(attrs["size"])
            if attrs["type"] == "integer":
                arr = np.frombuffer(payload, dtype="<i%s" % attrs["kind"], count=size)
            else:
                arr = np.frombuffer(payload, dtype="<c%d" % (2 * int(attrs["kind"])), count=size)
            out[name] = arr.copy()
            i += 1
    n = int(out["DIMENSION"][0])
    ns = int(out["NUMBER_SHIFT"][0])
    return dict(DIMENSION=n, NUMBER_SHIFT=ns, LIST_SHIFT=out["LIST_SHIFT"],
                LINEAR_OPERATOR=out["LINEAR_OPERATOR"].reshape((n, n), order="F"),
                RIGHT_HAND_SIDE=out["RIGHT_HAND_SIDE"],
                INCORRECT_SOLUTION=out["INCORRECT_SOLUTION"].reshape((n, ns), order="F"))


def read_linear_problem(path: str) -> dict:
    opener = bz2.open if path.endswith(".bz2") else open
    with opener(path, "rb") as f:
        return decode_linear_problem(f.read())
