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


def svmc_bytes(lower, upper=0, point=(0, 0, 0), normal=(128, 128, 128)):
    """one split-voxel (svmc) record as the 8 bytes the reference reads per voxel:
    {lower label, upper label, reference point x, y, z, normal x, y, z}
    (src/mcx_const.h:57)"""
    return np.array([lower, upper, *point, *normal], dtype=np.uint8)


def split_sphere_volume(dim, center, radius, inside, outside):
    """[8, dim, dim, dim] uint8 split-voxel volume of a ball (label inside) in a
    cube (label outside).  A voxel whose centre lies within 0.9 of the sphere
    and that holds the sphere point on the ray from the centre through it is
    split by the tangent plane there: reference point and normal quantised to
    bytes (p = byte/255 within the voxel, n = 2 byte/255 - 1).  The side with
    p.n above the plane offset carries the upper label, so the normal points
    into it: outward when the outside label is non-zero, inward otherwise
    (upper label 0 would mean an unsplit voxel)."""
    c = np.asarray(center, np.float64)
    idx = np.indices((dim, dim, dim)).astype(np.float64)
    q = np.moveaxis(idx, 0, -1) + 0.5
    d = np.linalg.norm(q - c, axis=-1)
    vol = np.zeros((8, dim, dim, dim), np.uint8)
    vol[0] = np.where(d < radius, inside, outside)
    near = np.abs(d - radius) < 0.9
    nhat = (q - c) / np.maximum(d, 1e-9)[..., None]
    s = c + radius * nhat
    off = s - np.moveaxis(idx, 0, -1)
    ok = near & np.all((off >= 0) & (off <= 1), axis=-1)
    outward = outside != 0
    upper, lower = (outside, inside) if outward else (inside, outside)
    nst = nhat if outward else -nhat
    vol[0][ok] = lower
    vol[1][ok] = upper
    pb = np.clip(np.rint(off * 255.0), 0, 255).astype(np.uint8)
    nb = np.clip(np.rint((nst + 1.0) * 127.5), 0, 255).astype(np.uint8)
    for k in range(3):
        vol[2 + k][ok] = pb[..., k][ok]
        vol[5 + k][ok] = nb[..., k][ok]
    return vol
