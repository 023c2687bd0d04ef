"""Small helpers shared by the tests (fixture decoding, hashing)."""
import hashlib

import numpy as np


def words(hexlist):
    return np.array([int(h, 16) for h in hexlist], dtype=np.uint64)


def f32(hexlist):
    return np.array([int(h, 16) for h in hexlist], dtype=np.uint32).view(np.float32)


def f64(hexlist):
    return np.array([int(h, 16) for h in hexlist], dtype=np.uint64).view(np.float64)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def lane_counts(a):
    """per-lane number of set bits over an array of uint64 words (numpy, vectorised)."""
    b = np.unpackbits(np.ascontiguousarray(a, dtype="<u8").view(np.uint8).reshape(-1, 8), axis=1,
                      bitorder="little")
    return b.sum(axis=0).astype(np.int64)
