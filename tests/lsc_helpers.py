"""Parameter sets and synthetic volumes shared by the test modules."""
import numpy as np

MODES = {
    "lossless": ((1, 1, 1, 1), 1.0),
    "b3-5-10-15_p0.5": ((3, 5, 10, 15), 0.5),
    "b3x4_p0.5": ((3, 3, 3, 3), 0.5),
    "b5x4_p0.5": ((5, 5, 5, 5), 0.5),
    "b10x4_p0.5": ((10, 10, 10, 10), 0.5),
    "b3-4-5-6_p0.7": ((3, 4, 5, 6), 0.7),
    "b1-2-4-8_p0": ((1, 2, 4, 8), 0.0),
    "b2-65535-70000-100000_p0.3": ((2, 65535, 70000, 100000), 0.3),
}


def synth(kind, dims, seed):
    """Test volume: datagen kinds plus a few adversarial ones."""
    from lscomp import datagen
    x, y, z = dims
    if kind in ("sparse", "cssi", "xpcs"):
        return datagen.volume(kind, seed, dims)
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.integers(0, 65536, size=dims, dtype=np.uint16)
    if kind == "zeros":
        return np.zeros(dims, np.uint16)
    if kind == "max":
        return np.full(dims, 65535, np.uint16)
    if kind == "mixed":
        v = rng.poisson(0.3, size=dims).astype(np.uint16)
        hot = rng.random(dims) < 0.01
        v[hot] = rng.integers(0, 65536, size=int(hot.sum()), dtype=np.uint16)
        return v
    raise ValueError(kind)
