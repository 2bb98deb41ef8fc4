"""A small FITS reader for the tests, written apart from the library's writer:
HDUs as (header bytes, data bytes), header cards as a dict, image cubes as
numpy arrays, and the FITS checksum convention (the 32-bit ones'-complement
sum of an HDU is -0 when its CHECKSUM card is right; DATASUM is the sum of
the data blocks).  write_bintables makes test inputs for the library's
table readers."""
from __future__ import annotations

import numpy as np

BLOCK = 2880


def split_hdus(buf: bytes) -> list[tuple[bytes, bytes]]:
    pos, out = 0, []
    while pos < len(buf):
        start = pos
        while True:
            block = buf[pos:pos + BLOCK]
            if len(block) < BLOCK:
                raise ValueError("truncated header")
            pos += BLOCK
            if any(block[i:i + 8] == b"END     " for i in range(0, BLOCK, 80)):
                break
        header = buf[start:pos]
        h = cards(header)
        naxis = int(h.get("NAXIS", 0))
        n = abs(int(h["BITPIX"])) // 8 if naxis else 0
        for k in range(naxis):
            n *= int(h[f"NAXIS{k + 1}"])
        n += int(h.get("PCOUNT", 0))
        n = (n + BLOCK - 1) // BLOCK * BLOCK
        out.append((header, buf[pos:pos + n]))
        pos += n
    return out


def _value(v: str):
    v = v.strip()
    if v.startswith("'"):
        end = 1
        while True:
            end = v.index("'", end)
            if end + 1 < len(v) and v[end + 1] == "'":
                end += 2
                continue
            return v[1:end].replace("''", "'").rstrip()
    v = v.split("/")[0].strip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v)


def cards(header: bytes) -> dict:
    out = {}
    for i in range(0, len(header), 80):
        c = header[i:i + 80].decode("ascii")
        if c.startswith("END") and c[3:].strip() == "":
            break
        if c[8:10] == "= ":
            out[c[:8].strip()] = _value(c[10:])
    return out


def card_list(header: bytes) -> list[str]:
    return [header[i:i + 80].decode("ascii") for i in range(0, len(header), 80)]


def image(header: bytes, data: bytes) -> np.ndarray:
    h = cards(header)
    dt = {-64: ">f8", -32: ">f4", 16: ">i2", 32: ">i4", 8: "u1"}[int(h["BITPIX"])]
    shape = [int(h[f"NAXIS{k}"]) for k in range(int(h["NAXIS"]), 0, -1)]
    n = int(np.prod(shape))
    return np.frombuffer(data, dtype=dt, count=n).reshape(shape)


def ones_sum(b: bytes, start: int = 0) -> int:
    w = np.frombuffer(b, dtype=">u4").astype(np.uint64)
    s = int(w.sum()) + start
    while s >> 32:
        s = (s & 0xFFFFFFFF) + (s >> 32)
    return s


def encode_checksum(hdu_sum: int) -> str:
    """The 16-character CHECKSUM value that brings hdu_sum to -0."""
    value = ~hdu_sum & 0xFFFFFFFF
    excluded = set(range(0x3A, 0x41)) | set(range(0x5B, 0x61))
    asc = [0] * 16
    for i in range(4):
        byte = (value >> (24 - 8 * i)) & 0xFF
        q, r = byte // 4 + 0x30, byte % 4
        ch = [q + r, q, q, q]
        moved = True
        while moved:
            moved = False
            for j in (0, 2):
                while ch[j] in excluded or ch[j + 1] in excluded:
                    ch[j] += 1
                    ch[j + 1] -= 1
                    moved = True
        for j in range(4):
            asc[4 * j + i] = ch[j]
    return "".join(chr(asc[(i + 15) % 16]) for i in range(16))


def verify(header: bytes, data: bytes) -> dict:
    """HDU and data sums, and the checksum re-encoded from a zeroed card."""
    h = cards(header)
    datasum = ones_sum(data)
    total = ones_sum(header, datasum)
    lst = card_list(header)
    k = next(i for i, c in enumerate(lst) if c.startswith("CHECKSUM"))
    zeroed = lst[k][:11] + "0" * 16 + lst[k][27:]
    hdr0 = "".join(lst[:k] + [zeroed] + lst[k + 1:]).encode("ascii")
    return {"hdu_sum_is_minus_zero": total == 0xFFFFFFFF,
            "datasum_ok": str(h.get("DATASUM")) == str(datasum),
            "reencoded": encode_checksum(ones_sum(hdr0, datasum)),
            "checksum": h.get("CHECKSUM")}


def _card(key: str, value) -> str:
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20}"
    elif isinstance(value, float):
        v = f"{repr(value):>20}"
    else:
        q = str(value).replace("'", "''")
        v = "'" + q.ljust(8) + "'"
    return f"{key:<8}= {v}".ljust(80)


def write_bintables(path, tables: list[tuple[str, list[tuple[str, np.ndarray]], dict]]):
    """A FITS file of binary tables: (EXTNAME, [(TTYPE, array [rows, repeat]
    of float32 / float64 / int32)], extra keywords)."""
    fmt = {np.dtype(np.float32): ("E", ">f4"), np.dtype(np.float64): ("D", ">f8"),
           np.dtype(np.int32): ("J", ">i4")}
    out = b"".join(_card(k, v).encode() for k, v in
                   (("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 0), ("EXTEND", True)))
    out += "END".ljust(80).encode()
    out += b" " * (-len(out) % BLOCK)
    for name, cols, keys in tables:
        nrows = cols[0][1].shape[0]
        row = []
        cards = [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 0),
                 ("NAXIS2", nrows), ("PCOUNT", 0), ("GCOUNT", 1), ("TFIELDS", len(cols))]
        width = 0
        for i, (ttype, a) in enumerate(cols, 1):
            a = np.asarray(a).reshape(nrows, -1)
            letter, be = fmt[a.dtype]
            cards += [(f"TTYPE{i}", ttype), (f"TFORM{i}", f"{a.shape[1]}{letter}")]
            row.append(a.astype(be))
            width += a.shape[1] * a.dtype.itemsize
        cards[3] = ("NAXIS1", width)
        cards += [("EXTNAME", name)] + list(keys.items())
        hdr = "".join(_card(k, v) for k, v in cards) + "END".ljust(80)
        hdr = hdr.encode() + b" " * (-len(hdr) % BLOCK)
        data = b"".join(b"".join(r[i].tobytes() for r in row) for i in range(nrows))
        out += hdr + data + b"\0" * (-len(data) % BLOCK)
    open(path, "wb").write(out)
