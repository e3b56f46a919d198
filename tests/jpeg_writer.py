"""A minimal baseline JPEG writer for the decoder tests, from ITU-T T.81:
any integral sampling factors, component ids, JFIF or Adobe markers, a
restart interval, and one interleaved scan or one scan per component. The
Huffman codes are fixed-length canonical codes over the symbols a scan uses
(valid, not compact). Only the tests use it: the bytes it writes are
decoded by both Pillow (libjpeg-turbo) and the scoring command's decoder."""
from __future__ import annotations

import math
import struct

import numpy as np

# Natural position of the k-th zigzag coefficient (T.81 figure A.6).
ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,
          7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
          39, 46, 53, 60, 61, 54, 47, 55, 62, 63]

_C = np.array([[(math.sqrt(0.125) if k == 0 else 0.5) * math.cos((2 * n + 1) * k * math.pi / 16)
                for n in range(8)] for k in range(8)])


def quant_table(scale: int) -> np.ndarray:
    """Natural-order quantisers growing with frequency."""
    u, v = np.meshgrid(np.arange(8), np.arange(8))
    return np.clip(1 + scale * (1 + u + v) // 2, 1, 255).astype(np.int64).ravel()


def _blocks(plane: np.ndarray, bw: int, bh: int, q: np.ndarray) -> np.ndarray:
    """Quantised DCT coefficients (natural order) of a plane padded by edge
    replication to bw x bh blocks."""
    h, w = plane.shape
    pad = np.pad(plane.astype(np.float64), ((0, bh * 8 - h), (0, bw * 8 - w)), mode="edge") - 128.0
    out = np.zeros((bh, bw, 64), np.int64)
    for by in range(bh):
        for bx in range(bw):
            f = _C @ pad[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] @ _C.T
            out[by, bx] = np.clip(np.round(f.ravel() / q), -1023, 1023)
    return out


def _category(v: int) -> int:
    return 0 if v == 0 else int(abs(v)).bit_length()


def _bits(v: int, s: int) -> int:
    return v if v >= 0 else v + (1 << s) - 1


class _BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, value: int, length: int):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc = self.n = 0

    def flush(self):
        while self.n:
            self.put(1, 1)


def _table(symbols) -> tuple[bytes, dict]:
    """DHT body counts + values and the code map of a fixed-length table."""
    used = sorted(set(symbols)) or [0]
    length = max(1, (len(used)).bit_length())  # len(used) < 2^length: no all-ones code
    counts = [0] * 16
    counts[length - 1] = len(used)
    return bytes(counts) + bytes(used), {s: (i, length) for i, s in enumerate(used)}


def _marker(code: int, body: bytes) -> bytes:
    return bytes([0xFF, code]) + struct.pack(">H", len(body) + 2) + body


def encode(planes, width: int, height: int, sampling, ids=None, q_scale: int = 4, jfif: bool = True,
           adobe_transform=None, restart: int = 0, separate_scans: bool = False) -> bytes:
    """`planes[i]` is component i at its downsampled size
    (ceil(width h_i / hmax) x ceil(height v_i / vmax)), uint8."""
    nc = len(planes)
    ids = ids or list(range(1, nc + 1))
    hmax = max(h for h, _ in sampling)
    vmax = max(v for _, v in sampling)
    mx = -(-width // (8 * hmax))
    my = -(-height // (8 * vmax))
    q = quant_table(q_scale)
    coefs = []
    for (h, v), p in zip(sampling, planes):
        dw, dh = -(-width * h // hmax), -(-height * v // vmax)
        assert p.shape == (dh, dw), (p.shape, dh, dw)
        coefs.append(_blocks(p, mx * h, my * v, q))

    out = bytearray(b"\xff\xd8")
    if jfif:
        out += _marker(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00")
    if adobe_transform is not None:
        out += _marker(0xEE, b"Adobe\x00\x64\x00\x00\x00\x00" + bytes([adobe_transform]))
    out += _marker(0xDB, bytes([0]) + bytes(int(q[ZIGZAG[k]]) for k in range(64)))
    sof = struct.pack(">BHHB", 8, height, width, nc)
    for i, (h, v) in enumerate(sampling):
        sof += bytes([ids[i], (h << 4) | v, 0])
    out += _marker(0xC0, sof)
    if restart:
        out += _marker(0xDD, struct.pack(">H", restart))

    scans = [[i] for i in range(nc)] if separate_scans else [list(range(nc))]
    for comps in scans:
        # The scan's blocks in decoding order, then its symbols.
        units = []
        if len(comps) == 1:
            c = comps[0]
            h, v = sampling[c]
            bw = -(-(-(-width * h // hmax)) // 8)
            bh = -(-(-(-height * v // vmax)) // 8)
            for y in range(bh):
                for x in range(bw):
                    units.append([(c, coefs[c][y, x])])
        else:
            for y in range(my):
                for x in range(mx):
                    mcu = []
                    for c in comps:
                        h, v = sampling[c]
                        for by in range(v):
                            for bx in range(h):
                                mcu.append((c, coefs[c][y * v + by, x * h + bx]))
                    units.append(mcu)
        events = []  # (kind, table component slot, symbol, extra value, extra length) or restart
        pred = {c: 0 for c in comps}
        for m, mcu in enumerate(units):
            if restart and m and m % restart == 0:
                events.append(("rst", (m // restart - 1) & 7))
                pred = {c: 0 for c in comps}
            for c, blk in mcu:
                zz = [int(blk[ZIGZAG[k]]) for k in range(64)]
                diff = zz[0] - pred[c]
                pred[c] = zz[0]
                s = _category(diff)
                events.append(("dc", c, s, _bits(diff, s), s))
                run = 0
                last = max([k for k in range(1, 64) if zz[k]] or [0])
                for k in range(1, last + 1):
                    if zz[k] == 0:
                        run += 1
                        continue
                    while run > 15:
                        events.append(("ac", c, 0xF0, 0, 0))
                        run -= 16
                    s = _category(zz[k])
                    events.append(("ac", c, (run << 4) | s, _bits(zz[k], s), s))
                    run = 0
                if last < 63:
                    events.append(("ac", c, 0x00, 0, 0))
        # Baseline allows two tables per class: the scan's first component
        # uses slot 0, the others share slot 1.
        slot = {c: min(i, 1) for i, c in enumerate(comps)}
        tables = {}
        dht = b""
        for kind, cls in (("dc", 0), ("ac", 1)):
            for t in sorted(set(slot.values())):
                body, codes = _table(e[2] for e in events if e[0] == kind and slot[e[1]] == t)
                tables[(kind, t)] = codes
                dht += bytes([(cls << 4) | t]) + body
        out += _marker(0xC4, dht)
        sos = bytes([len(comps)])
        for c in comps:
            sos += bytes([ids[c], (slot[c] << 4) | slot[c]])
        out += _marker(0xDA, sos + bytes([0, 63, 0]))
        bw = _BitWriter()
        for e in events:
            if e[0] == "rst":
                bw.flush()
                bw.out += bytes([0xFF, 0xD0 + e[1]])
                continue
            code, length = tables[(e[0], slot[e[1]])][e[2]]
            bw.put(code, length)
            if e[4]:
                bw.put(e[3], e[4])
        bw.flush()
        out += bw.out
    out += b"\xff\xd9"
    return bytes(out)
