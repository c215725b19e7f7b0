"""Pins for the CPU oracle: checks that fix its behaviour independently of its code.

Each ``pin_*`` takes an ``oracle.Oracle`` and raises AssertionError when the
oracle disagrees with:
  * hand-derived known-answer bytes (single 8x8 frames worked out from the
    reference kernel's arithmetic, lsCOMP/src/cuLSZ_kernel.cu:898-1277);
  * the bit layout implied by the reference's shifts (one-hot inputs: exactly
    one payload bit, at plane i, byte r, bit 7-k; pooled: byte t//8, bit 7-t%8);
  * decision boundaries computed by hand (pooling threshold in float,
    progressive bin choice, rate = bit width);
  * invariants (lossless round trip, lossy error bounds, stream size and
    offsets, header locality, frame-prefix property, serial == OpenMP bytes,
    block gather == numpy reshape of the zero-padded volume).
The same pins are run against oracle builds with one planted mistake each;
every mutant must fail at least one pin.
"""
import numpy as np

import oracle

LOSSLESS = (1, 1, 1, 1)


def _stream(orc, vol, bins, p):
    st, s, n = orc.compress(np.asarray(vol, np.uint16), bins, p)
    assert st == 0, f"compress status {st}"
    return bytes(s[:n])


# ---- hand-derived known answers ------------------------------------------

def pin_kat_bytes(orc):
    z = np.zeros((1, 8, 8), np.uint16)
    # all zeros, lossless: bc=3 (all counts 64), rate 0 -> 0x60, R = 4
    assert _stream(orc, z, LOSSLESS, 1.0) == bytes([0x60, 0, 0, 0])
    # v[0]=5, lossless: rate 3, planes 0..2 carry 1,0,1 in byte 0 bit 7
    v = z.copy(); v[0, 0, 0] = 5
    plane1 = bytes([0x80] + [0] * 7)
    assert _stream(orc, v, LOSSLESS, 1.0) == bytes([0x63, 0, 0, 0]) + plane1 + bytes(8) + plane1
    # v[0]=5, bins 3 5 10 15, p 0.5: zc=63 -> pool; counts (63,63,64,64) -> bc=1;
    # pair max floor(5/2)=2, 2/5=0 -> rate 0 -> 0x80|1<<5 = 0xA0
    assert _stream(orc, v, (3, 5, 10, 15), 0.5) == bytes([0xA0, 0, 0, 0])
    # v[0]=5, bins 1 1 1 1, p 0: pool (zc=63 > 0); qp[0] = 2 -> rate 2, hdr 0xE2;
    # plane 0 all zero, plane 1 byte 0 bit 7
    assert _stream(orc, v, LOSSLESS, 0.0) == bytes([0xE2, 0, 0, 0]) + bytes(4) + bytes([0x80, 0, 0, 0])
    # pair (1, 0) pooled: floor(1/2) = 0 -> rate 0 -> 0xE0; pair (3, 0): floor(3/2) = 1 -> 0xE1
    v1 = z.copy(); v1[0, 0, 0] = 1
    assert _stream(orc, v1, LOSSLESS, 0.0) == bytes([0xE0, 0, 0, 0])
    v3 = z.copy(); v3[0, 0, 0] = 3
    assert _stream(orc, v3, LOSSLESS, 0.0) == bytes([0xE1, 0, 0, 0, 0x80, 0, 0, 0])
    # all 7, bins 3 5 10 15, p 0.5: no zeros; counts (0,0,64,64) -> bc=1; 7/5=1 -> 0x21, one FF plane
    s = np.full((1, 8, 8), 7, np.uint16)
    assert _stream(orc, s, (3, 5, 10, 15), 0.5) == bytes([0x21, 0, 0, 0]) + b"\xff" * 8
    # all 65535 lossless: rate 16 -> 0x70, 16 planes of FF
    m = np.full((1, 8, 8), 65535, np.uint16)
    assert _stream(orc, m, LOSSLESS, 1.0) == bytes([0x70, 0, 0, 0]) + b"\xff" * 128
    # 1x8x9: two blocks, R = 4
    assert _stream(orc, np.zeros((1, 8, 9), np.uint16), LOSSLESS, 1.0) == bytes([0x60, 0x60, 0, 0])
    # 1x8x1: one block, the 7 missing columns are zero padding
    c = np.arange(1, 9, dtype=np.uint16).reshape(1, 8, 1)   # rows 0..7 hold 1..8 in column 0
    s = _stream(orc, c, LOSSLESS, 1.0)
    assert s[0] == 0x64 and len(s) == 4 + 32
    for i in range(4):   # plane i byte r bit 7 = bit i of (r+1)
        assert s[4 + 8 * i: 12 + 8 * i] == bytes([0x80 if ((r + 1) >> i) & 1 else 0 for r in range(8)])


def pin_kat_decode(orc):
    # the 0x21 block decodes to 1*5 everywhere; the 0xA0 block to zeros;
    # pooled 0xE2 block decodes both elements of pair 0 to 2
    assert np.all(orc.decode_block(0x21, b"\xff" * 8, (3, 5, 10, 15)) == 5)
    assert np.all(orc.decode_block(0xA0, b"", (3, 5, 10, 15)) == 0)
    d = orc.decode_block(0xE2, bytes(4) + bytes([0x80, 0, 0, 0]), LOSSLESS)
    assert d[0] == 2 and d[1] == 2 and np.all(d[2:] == 0)


# ---- bit layout -----------------------------------------------------------

def pin_onehot_plain(orc):
    for i in range(16):
        for r in range(8):
            for k in range(8):
                v = np.zeros(64, np.uint32)
                v[8 * r + k] = 1 << i
                h, nb = orc.block_header(v, LOSSLESS, 1.0)
                assert h == 0x60 | (i + 1) and nb == 8 * (i + 1), (i, r, k, h)
                pl = orc.encode_block(v, h, LOSSLESS)
                exp = bytearray(8 * (i + 1))
                exp[8 * i + r] = 0x80 >> k
                assert pl == bytes(exp), (i, r, k)
                assert np.array_equal(orc.decode_block(h, pl, LOSSLESS), v)


def pin_onehot_pooled(orc):
    for i in range(16):
        for t in range(32):
            r, c = divmod(t, 4)
            v = np.zeros(64, np.uint32)
            v[8 * r + 2 * c] = v[8 * r + 2 * c + 1] = 1 << i
            h, nb = orc.block_header(v, LOSSLESS, 0.0)   # zc = 62 > 0 -> pooled
            assert h == 0xE0 | (i + 1) and nb == 4 * (i + 1), (i, t, h)
            pl = orc.encode_block(v, h, LOSSLESS)
            exp = bytearray(4 * (i + 1))
            exp[4 * i + t // 8] = 0x80 >> (t % 8)
            assert pl == bytes(exp), (i, t)
            assert np.array_equal(orc.decode_block(h, pl, LOSSLESS), v)


# ---- decisions ------------------------------------------------------------

# smallest zero count that pools: (float)zc / 64 > p evaluated by hand
POOL_THRESHOLDS = {0.0: 1, 0.25: 17, 0.3: 20, 0.5: 33, 0.7: 45, 1.0: None}


def pin_pool_threshold(orc):
    for p, zmin in POOL_THRESHOLDS.items():
        for zc in range(0, 65):
            v = np.zeros(64, np.uint32)
            v[zc:] = 3
            h, _ = orc.block_header(v, LOSSLESS, p)
            want = zmin is not None and zc >= zmin
            assert bool(h & 0x80) == want, (p, zc, hex(h))


def pin_bin_choice(orc):
    bins = (3, 5, 10, 15)
    cases = [  # (nonzero values, expected bin choice)
        ([], 3), ([1, 2], 3), ([3], 0), ([4], 0), ([5], 1), ([9], 1), ([10], 2), ([14], 2),
        ([15], 3), ([60000], 3), ([3, 20], 0), ([7, 20], 1), ([12, 20], 2), ([2, 12], 2),
        ([4, 7, 12], 0), ([7, 12], 1),
    ]
    for vals, bc in cases:
        v = np.zeros(64, np.uint32)
        v[: len(vals)] = vals
        h, _ = orc.block_header(v, bins, 1.0)
        assert (h >> 5) & 3 == bc, (vals, hex(h))
        assert h & 0x1f == (max(vals, default=0) // bins[bc]).bit_length(), (vals, hex(h))
    # equal bins always choose 3; huge bins quantize everything to 0
    v = np.arange(64, dtype=np.uint32) * 1000
    assert (orc.block_header(v, (7, 7, 7, 7), 1.0)[0] >> 5) & 3 == 3
    assert orc.block_header(v, (70000, 80000, 90000, 100000), 1.0)[0] == 0x60


def pin_rate_edges(orc):
    for i in range(17):
        for val in {max(0, (1 << i) - 1), min(65535, 1 << i)}:
            v = np.zeros(64, np.uint32)
            v[17] = val
            h, nb = orc.block_header(v, LOSSLESS, 1.0)
            assert h & 0x1f == int(val).bit_length() and nb == 8 * int(val).bit_length(), (val, hex(h))


# ---- invariants -----------------------------------------------------------

def _volumes():
    rng = np.random.default_rng(1234)
    vols = []
    for shape in [(1, 8, 8), (2, 13, 17), (1, 9, 9), (3, 24, 40), (1, 8, 1), (2, 31, 66)]:
        vols.append(rng.poisson(0.4, size=shape).astype(np.uint16))
        u = rng.integers(0, 65536, size=shape).astype(np.uint16)
        u[rng.random(shape) < 0.5] = 0
        vols.append(u)
        vols.append((rng.integers(0, 2, size=shape) * rng.integers(0, 40, size=shape)).astype(np.uint16))
    return vols


def _blocks_numpy(vol):
    """(B, 64) blocks of a volume by zero padding and reshaping (independent of the oracle's indexing)."""
    x, y, z = vol.shape
    py, pz = -(-y // 8) * 8, -(-z // 8) * 8
    pad = np.zeros((x, py, pz), np.uint32)
    pad[:, :y, :z] = vol
    return pad.reshape(x, py // 8, 8, pz // 8, 8).transpose(0, 1, 3, 2, 4).reshape(-1, 64)


def _sizes(hdrs):
    hdrs = np.asarray(hdrs, np.int64)
    return (hdrs & 0x1f) * np.where(hdrs & 0x80, 4, 8)


def pin_lossless_roundtrip(orc):
    for vol in _volumes():
        st, s, n = orc.compress(vol, LOSSLESS, 1.0)
        st2, dec = orc.decompress(s[:n], vol.shape, LOSSLESS, fill=0xBEEF)
        assert st == 0 and st2 == 0 and np.array_equal(dec, vol)


def pin_stream_structure(orc):
    for vol in _volumes():
        for bins, p in ((LOSSLESS, 1.0), ((3, 5, 10, 15), 0.5), ((2, 2, 2, 2), 0.0)):
            st, s, n = orc.compress(vol, bins, p)
            blocks = _blocks_numpy(vol)
            nb = len(blocks)
            R = -(-nb // 4) * 4
            assert orc.num_blocks(vol.shape) == nb and orc.header_bytes(vol.shape) == R
            assert orc.max_compressed_bytes(vol.shape) == R + 128 * nb
            hdrs = np.frombuffer(bytes(s[:nb]), np.uint8)
            assert all(b == 0 for b in bytes(s[nb:R]))          # pad bytes
            sizes = _sizes(hdrs)
            assert n == R + int(sizes.sum())                  # size formula
            offs = R + np.concatenate([[0], np.cumsum(sizes)[:-1]])
            for b in range(nb):                               # headers from numpy-gathered blocks,
                h, nbytes = orc.block_header(blocks[b], bins, p)   # payload at the scanned offset
                assert h == hdrs[b] and nbytes == sizes[b]
                assert bytes(s[offs[b]: offs[b] + sizes[b]]) == orc.encode_block(blocks[b], h, bins)


def pin_lossy_bounds(orc):
    for vol in _volumes():
        for bins, p in (((3, 5, 10, 15), 0.5), ((3, 4, 5, 6), 0.7), ((2, 9, 9, 30), 0.0), ((5, 5, 5, 5), 1.0)):
            st, s, n = orc.compress(vol, bins, p)
            st2, dec = orc.decompress(s[:n], vol.shape, bins)
            blocks, dblocks = _blocks_numpy(vol).astype(np.int64), _blocks_numpy(dec).astype(np.int64)
            inside = _blocks_numpy(np.ones(vol.shape, np.uint16)).astype(bool)   # not padding
            hdrs = np.frombuffer(bytes(s[: len(blocks)]), np.uint8)
            for b in range(len(blocks)):
                h, v, w, m = int(hdrs[b]), blocks[b], dblocks[b], inside[b]
                bin_ = bins[(h >> 5) & 3]
                nz = int((v == 0).sum())                     # padding counts as zeros
                assert bool(h & 0x80) == (np.float32(nz) / np.float32(64) > np.float32(p))
                if h & 0x80:
                    a, c = v[0::2], v[1::2]
                    both = m[0::2] & m[1::2]
                    assert np.array_equal(w[0::2][both], w[1::2][both])
                    bound = (np.abs(a - c) + 1) // 2 + bin_ - 1
                    assert np.all((np.abs(a - w[0::2]) <= bound)[m[0::2]])
                    assert np.all((np.abs(c - w[1::2]) <= bound)[m[1::2]])
                else:
                    d = (v - w)[m]
                    assert np.all(d >= 0) and np.all(d <= bin_ - 1), (b, hex(h))


def pin_header_locality(orc):
    rng = np.random.default_rng(5)
    vol = rng.poisson(0.5, size=(2, 24, 24)).astype(np.uint16)
    st, s, n = orc.compress(vol, (3, 5, 10, 15), 0.5)
    vol2 = vol.copy()
    vol2[1, 8:16, 8:16] = rng.integers(0, 65536, size=(8, 8))    # block 4 of frame 1 only
    st, s2, n2 = orc.compress(vol2, (3, 5, 10, 15), 0.5)
    nb = orc.num_blocks(vol.shape)
    j = 9 + 4
    h1, h2 = np.frombuffer(bytes(s[:nb]), np.uint8), np.frombuffer(bytes(s2[:nb]), np.uint8)
    assert np.array_equal(np.delete(h1, j), np.delete(h2, j))
    R = orc.header_bytes(vol.shape)
    pre = R + int(_sizes(h1[:j]).sum())
    assert bytes(s[R:pre]) == bytes(s2[R:pre])
    tail1 = bytes(s[pre + int(_sizes(h1[j:j + 1]).sum()):n])
    tail2 = bytes(s2[pre + int(_sizes(h2[j:j + 1]).sum()):n2])
    assert tail1 == tail2


def pin_frame_prefix(orc):
    rng = np.random.default_rng(6)
    vol = rng.poisson(0.3, size=(5, 20, 28)).astype(np.uint16)
    st, s, n = orc.compress(vol, (3, 5, 10, 15), 0.5)
    nb, R = orc.num_blocks(vol.shape), orc.header_bytes(vol.shape)
    for k in (1, 2, 4):
        st, sk, nk = orc.compress(vol[:k], (3, 5, 10, 15), 0.5)
        nbk, Rk = orc.num_blocks(vol[:k].shape), orc.header_bytes(vol[:k].shape)
        assert bytes(sk[:nbk]) == bytes(s[:nbk])
        assert bytes(sk[Rk:nk]) == bytes(s[R: R + nk - Rk])


def pin_omp_equals_serial(orc):
    for vol in _volumes()[::3]:
        for bins, p in ((LOSSLESS, 1.0), ((3, 5, 10, 15), 0.5)):
            a = orc.compress(vol, bins, p)
            b = orc.compress(vol, bins, p, omp=True)
            assert a[0] == b[0] == 0 and a[2] == b[2] and np.array_equal(a[1], b[1])
            assert np.array_equal(orc.decompress(a[1], vol.shape, bins)[1],
                                  orc.decompress(a[1], vol.shape, bins, omp=True)[1])


# ---- uint32 volumes (cuLSZ_kernel.cu:4-824) ---------------------------------

def _u32_block(entries):
    v = np.zeros((1, 8, 8), np.uint32)
    for (r, k), val in entries.items():
        v[0, r, k] = val
    return v


def pin_u32_kat(orc):
    """Hand-derived uint32 streams: the wrapping pair average, rates above 16 and the rate-32 rejection."""
    # (0xFFFFFFFF + 3) mod 2^32 = 2 -> average 1; 62 zeros > 0.5*64 -> pooled, all bins 1 -> choice 3, rate 1
    st, s, n = orc.compress_u32(_u32_block({(0, 0): 0xFFFFFFFF, (0, 1): 3}), LOSSLESS, 0.5)
    assert st == 0 and bytes(s[:n]) == bytes([0xE1, 0, 0, 0, 0x80, 0, 0, 0])
    st, dec = orc.decompress_u32(s[:n], (1, 8, 8), LOSSLESS)
    assert st == 0 and dec[0, 0, 0] == 1 and dec[0, 0, 1] == 1 and not dec.ravel()[2:].any()
    # rate 20 plain: q[0][0] = 2^20 - 1 sets bit 7 of byte 0 in planes 0..19
    st, s, n = orc.compress_u32(_u32_block({(0, 0): (1 << 20) - 1}), LOSSLESS, 1.0)
    assert st == 0 and bytes(s[:n]) == bytes([0x74, 0, 0, 0]) + bytes([0x80, 0, 0, 0, 0, 0, 0, 0]) * 20
    # rate 31 plain: q[7][7] = 2^31 - 1 sets bit 0 of byte 7 in planes 0..30
    st, s, n = orc.compress_u32(_u32_block({(7, 7): (1 << 31) - 1}), LOSSLESS, 1.0)
    assert st == 0 and bytes(s[:n]) == bytes([0x7F, 0, 0, 0]) + bytes([0, 0, 0, 0, 0, 0, 0, 1]) * 31
    st, dec = orc.decompress_u32(s[:n], (1, 8, 8), LOSSLESS)
    assert st == 0 and dec[0, 7, 7] == (1 << 31) - 1
    # a quantized maximum of 2^31 needs rate 32, which the 5-bit field cannot hold: rejected;
    # with bin 2 the same block has rate 31
    v = _u32_block({(0, 5): 1 << 31})
    assert orc.compress_u32(v, LOSSLESS, 1.0)[0] == oracle.ERR_RANGE
    st, s, n = orc.compress_u32(v, (2, 2, 2, 2), 1.0)
    assert st == 0 and s[0] == 0x7F and n == 4 + 31 * 8


def _volumes_u32():
    rng = np.random.default_rng(4321)
    vols = []
    for shape in [(1, 8, 8), (2, 13, 17), (3, 24, 40), (1, 8, 1), (2, 31, 66)]:
        u = rng.integers(0, 1 << 31, size=shape, dtype=np.uint64).astype(np.uint32)
        u[rng.random(shape) < 0.5] = 0
        vols.append(u)
        vols.append((rng.poisson(2.0, size=shape) * rng.integers(1, 1 << 20, size=shape)).astype(np.uint32))
    return vols


def pin_u32_structure(orc):
    """Lossless round trip and the stream layout from numpy-gathered blocks, for uint32 volumes."""
    for vol in _volumes_u32():
        st, s, n = orc.compress_u32(vol, LOSSLESS, 1.0)
        st2, dec = orc.decompress_u32(s[:n], vol.shape, LOSSLESS, fill=0xDEADBEEF)
        assert st == 0 and st2 == 0 and np.array_equal(dec, vol)
        for bins, p in ((LOSSLESS, 1.0), ((3, 5, 10, 15), 0.5), ((1000, 1000, 70000, 70000), 0.25)):
            st, s, n = orc.compress_u32(vol, bins, p)
            blocks = _blocks_numpy(vol)
            nb = len(blocks)
            R = -(-nb // 4) * 4
            assert st == 0 and orc.max_compressed_bytes_u32(vol.shape) == R + 248 * nb
            hdrs = np.frombuffer(bytes(s[:nb]), np.uint8)
            assert all(b == 0 for b in bytes(s[nb:R]))
            sizes = _sizes(hdrs)
            assert n == R + int(sizes.sum())
            offs = R + np.concatenate([[0], np.cumsum(sizes)[:-1]])
            for b in range(nb):
                h, nbytes = orc.block_header(blocks[b], bins, p)
                assert h == hdrs[b] and nbytes == sizes[b]
                assert bytes(s[offs[b]: offs[b] + sizes[b]]) == orc.encode_block(blocks[b], h, bins)
            st2, dec = orc.decompress_u32(s[:n], vol.shape, bins)
            # plain blocks: 0 <= v - v' < bin (no pooling case)
            if p >= 1.0:
                d = vol.astype(np.int64) - dec.astype(np.int64)
                assert st2 == 0 and np.all(d >= 0) and np.all(d < max(bins))


def pin_u32_matches_u16(orc):
    """The reference's uint16 and uint32 kernels differ only in offset types: a uint32 volume holding
    uint16 values gives the uint16 stream, byte for byte."""
    for vol in _volumes()[:9]:
        for bins, p in ((LOSSLESS, 1.0), ((3, 5, 10, 15), 0.5), ((2, 9, 9, 30), 0.0)):
            a = orc.compress(vol, bins, p)
            b = orc.compress_u32(vol.astype(np.uint32), bins, p)
            assert a[0] == b[0] == 0 and a[2] == b[2] and np.array_equal(a[1], b[1])
            assert np.array_equal(orc.decompress(a[1], vol.shape, bins)[1].astype(np.uint32),
                                  orc.decompress_u32(b[1], vol.shape, bins)[1])


# ---- SSIM (cuLSZ_utility.cu:757-886) ----------------------------------------

def _ssim_numpy(a, b):
    """Mean SSIM over 7x7x7 windows at stride 2, vectorized with float64
    moments (an independent formulation of the reference's per-window loops)."""
    from numpy.lib.stride_tricks import sliding_window_view as swv
    x = swv(a.astype(np.float64), (7, 7, 7))[::2, ::2, ::2].reshape(-1, 343)
    y = swv(b.astype(np.float64), (7, 7, 7))[::2, ::2, ::2].reshape(-1, 343)
    mx, my = x.mean(1), y.mean(1)
    vx, vy = ((x - mx[:, None]) ** 2).mean(1), ((y - my[:, None]) ** 2).mean(1)
    cxy = ((x - mx[:, None]) * (y - my[:, None])).mean(1)
    r = x.max(1) - x.min(1)
    c1 = np.where(r == 0, 1e-4, 1e-4 * r * r)
    c2 = np.where(r == 0, 9e-4, 9e-4 * r * r)
    sx, sy = np.sqrt(vx), np.sqrt(vy)
    ssim = ((2 * mx * my + c1) / (mx * mx + my * my + c1) * (2 * sx * sy + c2) / (vx + vy + c2)
            * (cxy + c2 / 2) / (sx * sy + c2 / 2))
    return float(ssim.mean())


def pin_ssim(orc):
    from fractions import Fraction as F
    rng = np.random.default_rng(77)
    # one window, closed form: x = 10 everywhere, y = 10 but one sample 17.
    # Range 0 -> c1 = 1e-4, c2 = 9e-4; sigma_x = cov = 0 -> structure 1;
    # mean_y = 10 + 1/49, var_y = 49 * 342 / 343^2.
    x = np.full((7, 7, 7), 10, np.uint16)
    y = x.copy(); y[3, 2, 5] = 17
    my, vy = 10 + F(1, 49), F(49 * 342, 343 * 343)
    c1, c2 = F(1, 10 ** 4), F(9, 10 ** 4)
    want = float((20 * my + c1) / (100 + my * my + c1) * c2 / (vy + c2))
    assert abs(orc.ssim(x, y) - want) <= 1e-13 * want
    # identical volumes: 1
    v = rng.integers(0, 65536, (9, 12, 17)).astype(np.uint16)
    assert abs(orc.ssim(v, v) - 1.0) < 1e-13
    # no window fits: 0/0
    assert np.isnan(orc.ssim(v[:, :6], v[:, :6]))
    # against the vectorized formulation: uniform noise, sparse spikes (windows
    # whose original is constant), shapes with odd and even remainders
    cases = []
    a = rng.integers(0, 65536, (9, 11, 16)).astype(np.uint16)
    cases.append((a, np.clip(a.astype(np.int64) + rng.integers(-300, 301, a.shape), 0, 65535).astype(np.uint16)))
    s = np.zeros((8, 13, 20), np.uint16); s[rng.random(s.shape) < 0.02] = 900
    cases.append((s, (s // 3 * 3).astype(np.uint16)))
    c = rng.poisson(2.0, (10, 10, 10)).astype(np.uint16)
    cases.append((c, rng.poisson(2.0, c.shape).astype(np.uint16)))
    for a, b in cases:
        got, want = orc.ssim(a, b), _ssim_numpy(a, b)
        assert abs(got - want) <= 1e-12 * abs(want), (a.shape, got, want)


ALL_PINS = [pin_kat_bytes, pin_kat_decode, pin_onehot_plain, pin_onehot_pooled, pin_pool_threshold,
            pin_bin_choice, pin_rate_edges, pin_lossless_roundtrip, pin_stream_structure, pin_lossy_bounds,
            pin_header_locality, pin_frame_prefix, pin_omp_equals_serial, pin_u32_kat, pin_u32_structure,
            pin_u32_matches_u16, pin_ssim]

# planted mistakes in oracle/lscomp_oracle.c (ORACLE_MUTANT=n)
MUTANTS = {1: "plain bits LSB-first", 2: "plain block transposed", 3: "pool on >=", 4: "bin counts use <=",
           5: "no 4-byte header pad", 6: "pair average rounds up", 7: "rate one too high",
           8: "pooled decode drops the second pair element", 9: "SSIM windows moved by 1",
           10: "SSIM constants from the decoded range"}
