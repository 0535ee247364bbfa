"""The oracle's decoded picture hashes (src/hevc/hash.rs), pinned without
the reference binary: MD5 by the reference's own known answers
(hash.rs:255-261) and by Python's hashlib on every padding case; CRC-16 by
the published check value of the augmented CCITT CRC and by its closed form
over GF(2); the checksum by a vectorised restatement of its sum."""
import hashlib

import numpy as np
import pytest

import oracle


def test_md5_known_answers():
    assert oracle.md5(b"").hex() == "d41d8cd98f00b204e9800998ecf8427e"
    assert oracle.md5(b"abc").hex() == "900150983cd24fb0d6963f7d28e17f72"
    assert oracle.md5(b"The quick brown fox jumps over the lazy dog").hex() == "9e107d9d372bb6826bd81d3542a419d6"


@pytest.mark.parametrize("n", list(range(0, 140)) + [1000, 4095, 4096, 4097, 99841])
def test_md5_matches_hashlib(n):
    msg = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    assert oracle.md5(msg) == hashlib.md5(msg).digest()


def test_crc16_check_value():
    # CRC-16/AUG-CCITT: polynomial 0x1021, register 0xFFFF, message
    # augmented with 16 zero bits; the catalogue's check value.
    assert oracle.crc16(b"123456789") == 0xE5CC


def _crc_closed_form(msg: bytes) -> int:
    """x^16 (0xFFFF x^(8n) + M(x)) mod P over GF(2), P = x^16 + x^12 + x^5 + 1."""
    P = 0x11021

    def mod(v: int) -> int:
        while v.bit_length() > 16:
            v ^= P << (v.bit_length() - 17)
        return v

    m = int.from_bytes(msg, "big") if msg else 0
    return mod((0xFFFF << (8 * len(msg) + 16)) ^ (m << 16))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 64, 333])
def test_crc16_closed_form(n):
    msg = np.random.default_rng(100 + n).integers(0, 256, n, dtype=np.uint8).tobytes()
    assert oracle.crc16(msg) == _crc_closed_form(msg)


@pytest.mark.parametrize("w,h,wide", [(1, 1, False), (7, 5, False), (300, 260, False), (64, 33, True),
                                      (270, 257, True)])
def test_checksum_matches_vectorised_sum(w, h, wide):
    rng = np.random.default_rng(w * h + wide)
    plane = rng.integers(0, 1024 if wide else 256, (h, w), dtype=np.uint16 if wide else np.uint8)
    y, x = np.mgrid[0:h, 0:w]
    mask = ((x & 0xFF) ^ (y & 0xFF) ^ (x >> 8) ^ (y >> 8)).astype(np.uint64)
    if wide:
        want = int((((plane & 0xFF) ^ mask) + ((plane >> 8) ^ mask)).sum()) & 0xFFFFFFFF
    else:
        want = int((plane.astype(np.uint64) ^ mask).sum()) & 0xFFFFFFFF
    got = oracle.checksum(oracle.component_message(plane), w, wide)
    assert got == want


def test_message_is_low_byte_first():
    plane = np.array([[0x0102, 0x0304]], dtype=np.uint16)
    assert oracle.component_message(plane) == bytes([0x02, 0x01, 0x04, 0x03])
