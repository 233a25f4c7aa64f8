"""Helpers shared by the GPU parity tests."""
import numpy as np

from pipeline_amd import device as dev


def bits_of(words: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(np.ascontiguousarray(words, np.uint64).view(np.uint8), bitorder="little")[:n]


def mismatches(a: np.ndarray, b: np.ndarray, n: int) -> int:
    """Number of differing visibility bits among the first n, plus any set bit >= n."""
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    assert a.shape == b.shape
    diff = int(np.unpackbits((a ^ b).view(np.uint8)).sum())
    tail = bits_of(a, a.size * 64)[n:].sum() + bits_of(b, b.size * 64)[n:].sum()
    return diff + int(tail)


def gpu_cull_flat(lower, upper, matrices, vp, xform=None):
    scene = dev.upload_scene(lower, upper, matrices, xform)
    try:
        out = dev.cull(scene, vp)
        mask = out.download()
        out.free()
    finally:
        scene.free()
    return mask
