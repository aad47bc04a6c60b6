"""Oracle stages pinned by brute force and invariants where the fixtures (L=64 only) do not
reach: other lattice sizes, both codings, copy counts, scan and resampling."""
import math

import numpy as np
import pytest

import oracle

P = oracle._ptr
lib = oracle.lib()


def to_2d(sub, L):
    """Sublattice layout (gA[N/2] then gB[N/2]; A site (ty,tx) at x = 2tx + (ty&1)) -> L x L."""
    Nh = L * L // 2
    g = np.zeros((L, L), np.int64)
    for idx in range(Nh):
        ty, tx = divmod(idx, L // 2)
        g[ty, 2 * tx + (ty & 1)] = sub[idx]
        g[ty, 2 * tx + 1 - (ty & 1)] = sub[Nh + idx]
    return g


def from_2d(g):
    L = g.shape[0]
    Nh = L * L // 2
    sub = np.zeros(L * L, np.int8)
    for idx in range(Nh):
        ty, tx = divmod(idx, L // 2)
        sub[idx] = g[ty, 2 * tx + (ty & 1)]
        sub[Nh + idx] = g[ty, 2 * tx + 1 - (ty & 1)]
    return sub


DTYPES = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def random_words(rng, n, w):
    return rng.integers(0, 2**w, size=n, dtype=np.uint64).astype(DTYPES[w])


def energy_2d(g):
    return int(-(g * np.roll(g, -1, 1)).sum() - (g * np.roll(g, -1, 0)).sum()), int(g.sum())


@pytest.mark.parametrize("L", [4, 6, 8, 16, 24])
def test_ssc_energy_brute_force(L):
    rng = np.random.default_rng(L)
    R = 7
    sp = rng.choice(np.array([-1, 1], np.int8), size=(R, L * L))
    sp[0] = 1                          # ground state: E = -2N
    IE = np.zeros(R, np.int32)
    M = np.zeros(R, np.int32)
    lib.ora_ssc_energy(L, R, P(np.ascontiguousarray(sp)), P(IE), P(M))
    for r in range(R):
        e, m = energy_2d(to_2d(sp[r], L))
        assert (IE[r], M[r]) == (e, m)
    assert IE[0] == -2 * L * L and M[0] == L * L
    assert all(e % 4 == 0 and abs(e) <= 2 * L * L for e in IE)


@pytest.mark.parametrize("w", [8, 16, 32, 64])
@pytest.mark.parametrize("L", [4, 8, 24])
def test_msc_energy_equals_per_lane_ssc(w, L):
    rng = np.random.default_rng(w + L)
    rg = 3
    N = L * L
    words = random_words(rng, rg * N, w)
    IE = np.zeros(rg * w, np.int32)
    M = np.zeros(rg * w, np.int32)
    lib.ora_msc_energy(L, w, rg, P(words), P(IE), P(M))
    for g in range(rg):
        for b in range(w):
            bits = ((words[g * N:(g + 1) * N].astype(np.uint64) >> np.uint64(b)) & np.uint64(1)).astype(np.int8)
            spins = (2 * bits - 1).astype(np.int8)
            e, m = energy_2d(to_2d(spins, L))
            assert (IE[g * w + b], M[g * w + b]) == (e, m)


def _sweep_2d_ssc(g, L, sweeps, seed, sub, umap):
    """One replica's sweeps from 2-D coordinates: site (h, idx) uses word
    (2s+h)*n_t + idx//128 of stream sub + idx%128 (n_t = ceil((N/2 - t)/128))."""
    Nh = L * L // 2
    for s in range(sweeps):
        for h in range(2):
            new = g.copy()
            for idx in range(Nh):
                ty, tx = divmod(idx, L // 2)
                x = 2 * tx + ((ty & 1) if h == 0 else 1 - (ty & 1))
                t, k = idx % 128, idx // 128
                nt = (Nh - t + 127) // 128
                v = lib.ora_stream_word(seed, sub + t, (2 * s + h) * nt + k)
                u = lib.ora_uniform(v, 0, umap)
                nb = g[ty, (x + 1) % L] + g[ty, (x - 1) % L] + g[(ty + 1) % L, x] + g[(ty - 1) % L, x]
                dE = 2 * g[ty, x] * nb
                T = min(1.0, math.exp(-dE * beta_of[0])) if dE > 0 else 1.0
                if u < T:
                    new[ty, x] = -g[ty, x]
            g = new
    return g


beta_of = [0.0]


@pytest.mark.parametrize("L", [4, 6, 10, 18])
def test_ssc_sweep_geometry_and_streams_2d(L):
    beta = 0.7
    beta_of[0] = beta
    R, seed, IS = 2, 12345, 77
    rng = np.random.default_rng(L)
    sp = rng.choice(np.array([-1, 1], np.int8), size=(R, L * L))
    ref = []
    for r in range(R):
        g = _sweep_2d_ssc(to_2d(sp[r], L), L, 2, seed, IS + 128 * r, oracle.U32_OPEN_CLOSED)
        ref.append(from_2d(g))
    lib.ora_ssc_sweep(L, R, P(sp), 2, seed, IS, beta, oracle.U32_OPEN_CLOSED)
    for r in range(R):
        assert (sp[r] == ref[r]).all()


@pytest.mark.parametrize("L,w", [(4, 32), (6, 64), (10, 32), (4, 8), (6, 16), (10, 8)])
def test_msc_sweep_per_lane_2d(L, w):
    """Each replica bit follows the MSC rule: its random is LCG step `bit` from the site's
    stream word; count>=2 antiparallel flips, count 1 iff ran<T4, count 0 iff ran<T8."""
    beta, seed, IS, rg = 0.3, 99, 5, 1
    N, Nh = L * L, L * L // 2
    T4 = math.ceil(4294967296.0 * math.exp(-4 * beta))
    T8 = math.ceil(4294967296.0 * math.exp(-8 * beta))
    rng = np.random.default_rng(L * w)
    words = random_words(rng, rg * N, w)
    lanes = [to_2d((2 * ((words.astype(np.uint64) >> np.uint64(b)) & np.uint64(1)).astype(np.int64) - 1), L)
             for b in range(w)]
    for s in range(1):
        for h in range(2):
            news = [g.copy() for g in lanes]
            for idx in range(Nh):
                ty, tx = divmod(idx, L // 2)
                x = 2 * tx + ((ty & 1) if h == 0 else 1 - (ty & 1))
                t, k = idx % 128, idx // 128
                nt = (Nh - t + 127) // 128
                ran = lib.ora_stream_word(seed, IS + t, (2 * s + h) * nt + k)
                for b in range(w):
                    g = lanes[b]
                    anti = sum(int(g[ty, x] != n) for n in (g[ty, (x + 1) % L], g[ty, (x - 1) % L],
                                                             g[(ty + 1) % L, x], g[(ty - 1) % L, x]))
                    acc = anti >= 2 or (anti == 1 and ran < T4) or (anti == 0 and ran < T8)
                    if acc:
                        news[b][ty, x] = -g[ty, x]
                    ran = (1664525 * ran + 1013904223) & 0xFFFFFFFF
            lanes = news
    lib.ora_msc_sweep(L, w, rg, P(words), 1, seed, IS, beta)
    for b in range(w):
        got = (2 * ((words.astype(np.uint64) >> np.uint64(b)) & np.uint64(1)).astype(np.int64) - 1)
        assert (got == from_2d(lanes[b]).astype(np.int64)).all()


def test_copy_counts_expectation_and_scan():
    """E[Roff_i] = mu_i and sum mu_i = Rinit (nearest-integer resampling, SSC:243-247)."""
    R, Rinit = 20000, 20000
    rng = np.random.default_rng(5)
    IE = (4 * rng.integers(-300, 300, R)).astype(np.int32)
    dB, Emean = 0.01, float(IE.mean())
    q = lib.ora_ssc_q(R, P(IE), dB, Emean)
    lnQ = lib.ora_lnq(dB, Emean, q, R)
    mu = Rinit / R * np.exp(-dB * IE.astype(np.float64) - lnQ)
    assert abs(mu.sum() - Rinit) < 1e-6 * Rinit
    tot = []
    for IS in range(0, 40 * R, R):
        Roff = np.zeros(R, np.uint32)
        n = lib.ora_ssc_tau(R, P(IE), Rinit, dB, lnQ, 7, IS, oracle.U32_OPEN_CLOSED, P(Roff))
        assert n == int(Roff.sum())
        assert ((Roff == np.floor(mu)) | (Roff == np.floor(mu) + 1)).all()
        tot.append(Roff.astype(np.float64))
    emp = np.mean(tot, axis=0)
    frac = mu - np.floor(mu)
    se = np.sqrt(frac * (1 - frac) / len(tot)) + 1e-12
    assert np.mean(np.abs(emp - mu) / se) < 1.2


def test_resample_order_ssc():
    L, R = 4, 5
    sp = np.arange(R * 16, dtype=np.int64).astype(np.int8).reshape(R, 16)
    Roff = np.array([2, 0, 1, 3, 0], np.uint32)
    out = np.zeros((6, 16), np.int8)
    lib.ora_ssc_resample(L, R, P(sp), P(Roff), P(out))
    assert [int(x) for x in (out[:, 0])] == [sp[0, 0], sp[0, 0], sp[2, 0], sp[3, 0], sp[3, 0], sp[3, 0]]


@pytest.mark.parametrize("w", [8, 16, 32, 64])
def test_resample_order_msc_transposed(w):
    """Child n = prefix + p -> group n mod rg_new, lane n div rg_new (MSC:413-414); for 8 and
    16 bits the reference's atomicXor on the 32-bit word holding 4 / 2 sites sets bit
    it_k + 8 (t & 3) / it_k + 16 (t & 1) of it (MSC:415-422), i.e. bit it_k of site t."""
    L, rg_old = 4, 2
    N = 16
    rng = np.random.default_rng(9 + w)
    words = random_words(rng, rg_old * N, w)
    act = np.ones(rg_old * w, np.uint8)
    IE = np.arange(rg_old * w, dtype=np.int32)
    Roff = rng.integers(0, 3, size=rg_old * w).astype(np.uint32)
    Rn = int(Roff.sum())
    rg_new = math.ceil(Rn / w)
    wn = np.zeros(rg_new * N, DTYPES[w])
    an = np.zeros(rg_new * w, np.uint8)
    IEn = np.zeros(rg_new * w, np.int32)
    lib.ora_msc_resample(L, w, rg_old, P(words), P(act), P(IE), P(Roff), Rn, P(wn), P(an), P(IEn))
    parents = np.repeat(np.arange(rg_old * w), Roff)
    for n in range(Rn):
        j, k = n % rg_new, n // rg_new
        p = int(parents[n])
        pg, pb = divmod(p, w)
        for s in range(N):
            assert (int(wn[j * N + s]) >> k) & 1 == (int(words[pg * N + s]) >> pb) & 1
        assert an[j * w + k] == 1 and IEn[j * w + k] == IE[p]
    assert an.sum() == Rn


@pytest.mark.parametrize("w", [8, 16])
def test_msc_init_narrow_words_are_low_bits_of_one_stream_word(w):
    """MSbits 8 / 16: gA[idx] = curand() & 0xFF / 0xFFFF, one stream word per site word in the
    same order as 32 bits (MSC:173-178), so the words are the low bits of the 32-bit init's."""
    L, rg, R, seed, IS = 6, 3, 3 * w - 2, 41, 1000
    N = L * L
    ref = np.zeros(rg * N, np.uint32)
    act32 = np.zeros(rg * 32, np.uint8)
    lib.ora_msc_init(L, 32, rg, R, seed, IS, P(ref), P(act32))
    got = np.zeros(rg * N, DTYPES[w])
    act = np.zeros(rg * w, np.uint8)
    lib.ora_msc_init(L, w, rg, R, seed, IS, P(got), P(act))
    assert (got.astype(np.uint32) == (ref & np.uint32(2**w - 1))).all()
    assert act.sum() == R and (act.reshape(rg, w)[:-1] == 1).all()


@pytest.mark.parametrize("w", [8, 16])
def test_small_block_tree(w):
    """smallblockReduceSum (MSC:96-105) pairs shared[t] with shared[t + stride] for stride =
    w/2 .. 1; for 8 lanes that is ((v0+v4)+(v2+v6))+((v1+v5)+(v3+v7)).  It equals the
    32-lane shuffle-down tree of the 32/64-bit path on the same values padded with zeros
    (adding +0.0 is exact), which is how the device evaluates it."""
    rng = np.random.default_rng(w)
    for _ in range(200):
        v = np.exp(rng.normal(0, 30, size=w))
        got = lib.ora_small_block_reduce(P(v), w)
        if w == 8:
            assert got == ((v[0] + v[4]) + (v[2] + v[6])) + ((v[1] + v[5]) + (v[3] + v[7]))
        pad = np.zeros(32)
        pad[:w] = v
        assert got == lib.ora_block_reduce(P(pad), 32)
