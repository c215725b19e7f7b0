"""Synthetic volume generator: determinism, frame-prefix property, value statistics."""
import numpy as np
import pytest


@pytest.mark.parametrize("kind", ["sparse", "cssi", "xpcs"])
def test_deterministic_and_prefix(kind):
    from lscomp import datagen
    a = datagen.volume(kind, 7, (6, 64, 1100))
    b = datagen.volume(kind, 7, (6, 64, 1100))
    assert np.array_equal(a, b)
    assert np.array_equal(datagen.volume(kind, 7, (2, 64, 1100)), a[:2])
    assert np.array_equal(datagen.volume(kind, 7, (2, 64, 1100), first_frame=3), a[3:5])
    assert not np.array_equal(datagen.volume(kind, 8, (6, 64, 1100)), a)


def test_statistics():
    from lscomp import datagen
    s = datagen.volume("sparse", 1, (4, 256, 256))
    assert 0.94 < (s == 0).mean() < 0.96                      # Poisson(0.05) + 0.1% hot
    assert 0.0005 < (s >= 200).mean() < 0.0015
    c = datagen.volume("cssi", 2, (3, 1813, 1558))
    assert 0.72 < (c == 0).mean() < 0.76                      # Poisson(0.3) + gaps
    assert np.all(c[:, :, 512:520] == 0) and np.all(c[:, :, 1032:1040] == 0)
    assert c.max() > 10000                                   # Bragg peaks
    x = datagen.volume("xpcs", 3, (4, 512, 512))
    assert 0.015 < x.mean() < 0.025 and x.max() < 100       # mean 0.02 speckle
