"""The self-describing container (lscomp_container_pack / _parse): a round trip
of oracle-made uint16 and uint32 streams, and rejection of malformed containers.
Host memory only, so no GPU is needed."""
import struct

import numpy as np
import pytest

import lscomp
from lsc_helpers import synth


def _entries(orc):
    rng = np.random.default_rng(3)
    out = []
    for i, (shape, bins, p) in enumerate([((2, 13, 17), (1, 1, 1, 1), 1.0), ((1, 64, 64), (3, 5, 10, 15), 0.5),
                                          ((3, 9, 10), (2, 2, 2, 2), 0.0)]):
        vol = synth("cssi", shape, seed=i)
        st, s, n = orc.compress(vol, bins, p)
        out.append(({"stream": s[:n].tobytes(), "dims": shape, "elem_bytes": 2, "bins": bins, "pooling": p,
                     "tag": 100 + i}, vol))
    vol = rng.integers(0, 1 << 30, size=(2, 20, 28), dtype=np.uint64).astype(np.uint32)
    st, s, n = orc.compress_u32(vol, (1, 1, 1, 1), 1.0)
    out.append(({"stream": s[:n].tobytes(), "dims": vol.shape, "elem_bytes": 4, "bins": (1, 1, 1, 1),
                 "pooling": 1.0, "tag": 7}, vol))
    return out


def test_roundtrip(orc):
    ents = _entries(orc)
    buf = lscomp.container_pack([e for e, _ in ents])
    assert buf[:4] == b"LSCP" and struct.unpack_from("<I", buf, 8)[0] == len(ents)
    assert struct.unpack_from("<Q", buf, 16)[0] == len(buf)
    items = lscomp.container_parse(buf)
    assert len(items) == len(ents)
    for it, (e, vol) in zip(items, ents):
        assert it["offset"] % 16 == 0
        assert tuple(it["dims"]) == tuple(e["dims"]) and it["elem_bytes"] == e["elem_bytes"]
        assert tuple(it["bins"]) == tuple(e["bins"]) and it["tag"] == e["tag"]
        assert abs(it["pooling"] - e["pooling"]) < 1e-7
        assert bytes(it["stream"]) == e["stream"]
        if e["elem_bytes"] == 2:
            st, dec = orc.decompress(bytes(it["stream"]), it["dims"], it["bins"])
        else:
            st, dec = orc.decompress_u32(bytes(it["stream"]), it["dims"], it["bins"])
        assert st == 0
        if tuple(e["bins"]) == (1, 1, 1, 1) and e["pooling"] >= 1.0:
            assert np.array_equal(dec, vol)


def test_empty_container():
    buf = lscomp.container_pack([])
    assert len(buf) == 48 and lscomp.container_parse(buf) == []


def _corrupt(buf, off, fmt, value):
    b = bytearray(buf)
    struct.pack_into(fmt, b, off, value)
    return bytes(b)


def test_rejects_malformed(orc):
    ents = _entries(orc)
    buf = lscomp.container_pack([e for e, _ in ents])
    e1 = 48 + 64   # second index entry
    bad = [
        b"LSCX" + buf[4:],                               # magic
        _corrupt(buf, 4, "<H", 2),                       # version
        _corrupt(buf, 8, "<I", 1000),                    # more items than the container holds
        buf[: len(buf) - 16],                            # truncated (total bytes > length)
        _corrupt(buf, e1 + 40, "<Q", 8),                 # stream offset inside the index
        _corrupt(buf, e1 + 40, "<Q", len(buf)),          # stream past the end
        _corrupt(buf, e1 + 40, "<Q", 49),                # misaligned offset
        _corrupt(buf, e1 + 16, "<I", 7),                 # bins not ascending (7 > 5)
        _corrupt(buf, e1 + 12, "<I", 3),                 # element size
        _corrupt(buf, e1 + 48, "<Q", 3),                 # shorter than its header section
    ]
    for b in bad:
        with pytest.raises(lscomp.LscompError) as e:
            lscomp.container_parse(b)
        assert e.value.status == lscomp.ERR_CORRUPT


def test_pack_validates(orc):
    e, _ = _entries(orc)[1]
    for key, val in (("bins", (5, 3, 10, 15)), ("elem_bytes", 8), ("pooling", 1.5), ("dims", (0, 8, 8))):
        bad = dict(e, **{key: val})
        with pytest.raises(lscomp.LscompError) as err:
            lscomp.container_pack([bad])
        assert err.value.status == lscomp.ERR_INVALID_ARG
    with pytest.raises(lscomp.LscompError):   # stream larger than the worst case for its dims
        lscomp.container_pack([dict(e, stream=e["stream"] * 200)])
