"""Shared helpers for the test suite."""

import numpy as np


def sorted_records(detp):
    """Detected-photon records ([reclen, n] float32) as uint32 bit rows in lexicographic order.

    GPU records are appended in atomic order, so record sets are compared sorted."""
    bits = np.ascontiguousarray(np.asarray(detp, dtype=np.float32).T).view(np.uint32)
    return bits[np.lexsort(bits.T[::-1])]


def sorted_seeds(seeds):
    """[16, n] uint8 seed columns as rows in lexicographic order."""
    rows = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint8).T)
    return rows[np.lexsort(rows.T[::-1])]


def xorshift128p_floats(seed4, n):
    """Independent numpy statement of xorshift128+ (Vigna 2014, shifts 23/18/5) with the
    reference's float mapping: seeds t0=s0<<32|s1, t1=s2<<32|s3; output
    float32((low32(t1'+s0) >> 9) | 0x3F800000) - 1."""
    mask = (1 << 64) - 1
    t0 = (int(seed4[0]) << 32) | int(seed4[1])
    t1 = (int(seed4[2]) << 32) | int(seed4[3])
    out = np.zeros(n, dtype=np.float32)
    for i in range(n):
        s1 = t0
        s0 = t1
        t0 = s0
        s1 ^= (s1 << 23) & mask
        t1 = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5)
        low = ((t1 + s0) & mask) & 0xFFFFFFFF
        bits = np.uint32(0x3F800000 | (low >> 9))
        out[i] = bits.view(np.float32) - np.float32(1.0)
    return out
