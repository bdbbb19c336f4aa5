"""Integer and bit-level arithmetic of the tiled kernels, checked on the CPU.

The projector's halo tile (srf_amd/csrc/srf_tiled.hip, march_tile with
HALO): a cell index clamped into [origin - 1, origin + extent] by one med3
replaces the per-term tile test (unsigned)(c - origin) < extent.  Checked here
exhaustively on the CPU over origins, in-image extents and cell indices:

- a cell the test accepts is left unchanged by the clamp;
- a cell the test rejects is moved to origin - 1 or origin + extent, and that
  cell is a zero: a halo cell (-1 or TX after subtracting the origin) or a tile
  cell beyond the image edge, both staged as 0;
- every clamped index lies inside the halo tile's range [-1, TX], so its LDS
  word lies inside the tile;
- the byte address formed from it by signed 24-bit multiply-adds
  (v_mad_i32_i24) is that word's, also for tiles at the image origin where the
  clamped index is -1.

The backprojector's fixed-point terms: one double FMA onto 1.5 * 2^52 gives
the rounded integer product.
"""
import numpy as np

TX, TZ = 32, 8


def med3(x, lo, hi):
    return np.minimum(np.maximum(x, lo), hi)


def staged_zero(local, extent, t):
    """The staging kernel writes 0 at local index -1 and t (the halo) and at
    local indices >= extent (past the image edge)."""
    return (local == -1) | (local == t) | (local >= extent)


def check_axis(t, n_image):
    for origin in range(0, n_image, t):
        extent = min(t, n_image - origin)
        c = np.arange(-40, n_image + 40)
        accepted = ((c - origin) >= 0) & ((c - origin) < extent)
        cl = med3(c, origin - 1, origin + extent)
        local = cl - origin
        assert np.array_equal(cl[accepted], c[accepted])
        assert np.all(staged_zero(local[~accepted], extent, t))
        assert local.min() >= -1 and local.max() <= t


def test_clamp_equals_tile_test_secondary():
    for n in (32, 33, 50, 64, 96, 200, 256):
        check_axis(TX, n)


def test_clamp_equals_tile_test_axial():
    for n in (8, 9, 13, 64, 128, 200):
        check_axis(TZ, n)


def test_halo_words_inside_tile():
    """Word (lx+1) + HSY (ly+1) + HSZ (lz+1) of every cell with -1 <= l <= T
    lies in [0, HPAD), and the dummy word HPAD is past all of them."""
    hx, hy, hz = TX + 2, TX + 2, TZ + 2
    hsy = hx + 1
    hsz = hsy * hy + 1
    hpad = hsz * (hz - 1) + hsy * (hy - 1) + hx
    assert hpad == 11908
    l = np.arange(-1, TX + 1)
    lz = np.arange(-1, TZ + 1)
    words = (l[:, None, None] + 1) + hsy * (l[None, :, None] + 1) + hsz * (lz[None, None, :] + 1)
    assert words.min() == 0 and words.max() == hpad - 1
    assert len(np.unique(words)) == words.size     # distinct cells, distinct words


def _wrap32(v):
    return ((np.asarray(v, np.int64) + 2 ** 31) % 2 ** 32) - 2 ** 31


def mad_i24(a, b, c):
    """v_mad_i32_i24: sign-extended low 24 bits of a and b, product plus c,
    low 32 bits."""
    def sext24(v):
        v = np.asarray(v, np.int64) & 0xFFFFFF
        return np.where(v >= 2 ** 23, v - 2 ** 24, v)
    return _wrap32(sext24(a) * sext24(b) + np.asarray(c, np.int64))


def mad_u24(a, b, c):
    """v_mad_u32_u24: zero-extended low 24 bits of a and b."""
    return _wrap32((np.asarray(a, np.int64) & 0xFFFFFF) * (np.asarray(b, np.int64) & 0xFFFFFF)
                   + np.asarray(c, np.int64))


def halo_addresses(xdom, p0, s0, z0, es, ez, mad):
    """Byte addresses the projector's halo marcher forms for every slab of the
    tile's primary range and every clamped (secondary, axial) cell, as
    march_tile computes them: base folded with the tile origin, then two
    24-bit multiply-adds.  Returns (addresses, expected word * 4)."""
    hx, hy, hz = TX + 2, TX + 2, TZ + 2
    hsy = hx + 1
    hsz = hsy * hy + 1
    es_b = 4
    sp = 1 if xdom else hsy
    ss = hsy if xdom else 1
    bs, bz = ss * es_b, hsz * es_b
    lp = np.arange(0, TX)
    a = p0 + lp                                   # slab index
    base = (a - p0 + 1) * sp * es_b - (s0 - 1) * bs - (z0 - 1) * bz
    c = np.arange(s0 - 1, s0 + es + 1)            # clamp range along secondary
    d = np.arange(z0 - 1, z0 + ez + 1)            # and along z
    A = mad(c[None, :, None], bs, base[:, None, None])
    addr = mad(d[None, None, :], bz, A)
    word = (lp[:, None, None] + 1) * sp + (c[None, :, None] - s0 + 1) * ss + (d[None, None, :] - z0 + 1) * hsz
    return addr, word * 4


def test_halo_addresses_in_tile_at_image_origin():
    """Every address the halo marcher forms, for tiles at the image origin
    (s0 = 0, z0 = 0: the clamped index is -1) and elsewhere, is the intended
    halo word and lies inside the tile's LDS words.  The unsigned 24-bit form
    maps index -1 to 2^24 - 1 and leaves the tile."""
    hx, hy, hz = TX + 2, TX + 2, TZ + 2
    hsy = hx + 1
    hsz = hsy * hy + 1
    hpad = hsz * (hz - 1) + hsy * (hy - 1) + hx
    for xdom in (True, False):
        for n_s, n_z in ((256, 128), (96, 64), (50, 13)):
            for s0 in range(0, n_s, TX):
                for z0 in range(0, n_z, TZ):
                    es, ez = min(TX, n_s - s0), min(TZ, n_z - z0)
                    addr, want = halo_addresses(xdom, 64, s0, z0, es, ez, mad_i24)
                    assert np.array_equal(addr, want)
                    assert addr.min() >= 0 and addr.max() < hpad * 4
    addr, want = halo_addresses(True, 0, 0, 0, TX, TZ, mad_u24)
    assert addr.max() >= 2 ** 24 and not np.array_equal(addr, want)


def test_fixed_point_term_is_rounded_product():
    """The backprojector's fixed-point term (srf_tiled.hip, fixed_term): the bit
    pattern of fma(v, r, 1.5 * 2^52) minus 0x4338000000000000 is round(v * r)
    (ties to even) as a two's-complement int64, for |v * r| < 2^51.  v is a
    float32 value widened to double and r a float32 fraction, so v * r is exact
    in double and the fma equals the product plus the constant."""
    rng = np.random.default_rng(7)
    v = (rng.standard_normal(200000) * rng.choice([1.0, 1e3, 1e6, 2.0 ** 40], 200000)).astype(np.float32)
    r = rng.random(200000).astype(np.float32)
    prod = v.astype(np.float64) * r.astype(np.float64)
    keep = np.abs(prod) < 2.0 ** 51
    prod = prod[keep]
    s = prod + 6755399441055744.0
    got = s.view(np.int64) - np.int64(0x4338000000000000)
    assert np.array_equal(got, np.rint(prod).astype(np.int64))
    # ties go to even, negative values wrap as two's complement
    t = np.array([2.5, 3.5, -2.5, -0.5, 0.5, -1.0], np.float64)
    got_t = (t + 6755399441055744.0).view(np.int64) - np.int64(0x4338000000000000)
    assert got_t.tolist() == [2, 4, -2, 0, 0, -1]


def test_backprojection_value_bands():
    """The backprojector's exponent bands (srf_tiled.hip, bp_item /
    bp_item_bands): with vtop < 2^etop the item's largest value, a nonzero
    value v belongs to band min((etop - e_v) // 24, nb - 1) (e_v: |v| < 2^e_v)
    and band b's shift is 49 - ex(c_b), c_0 = vtop * W, c_b = 2^(etop - 24 b) W.
    Emulated on wide random items: every value lands in one band, no term of a
    band reaches 2^49 after its shift (no overflow), and outside the last band
    every value keeps at least 2^24 of resolution (a band spans 2^24)."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = int(rng.integers(1, 8193))
        v = (rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-30, 30, n)).astype(np.float32)
        v[rng.random(n) < 0.05] = 0.0
        wmax = np.float32(10.0 ** rng.uniform(-3, 2))
        nz = v[v != 0]
        if nz.size == 0:
            continue
        a = np.abs(nz).astype(np.float64)
        vtop, vbot = a.max(), a.min()
        etop, ebot = np.frexp(np.float32(vtop))[1], np.frexp(np.float32(vbot))[1]
        nb = min((etop - ebot) // 24 + 1, 8)
        ev = np.frexp(nz)[1]
        band = np.minimum((etop - ev) // 24, nb - 1)
        assert band.min() >= 0 and band.max() <= nb - 1
        for b in range(nb):
            c = np.float32(vtop) * wmax if b == 0 else np.ldexp(wmax, int(etop - 24 * b))
            shift = 49 - int(np.frexp(np.float32(c))[1])
            vb = a[band == b]
            if vb.size == 0:
                continue
            scaled = np.ldexp(vb * np.float64(wmax), shift)
            assert scaled.max() < 2.0 ** 49
            if b < nb - 1:
                assert scaled.min() >= 2.0 ** 24
