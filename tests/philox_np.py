"""numpy restatement of the device thermal source: Philox4x32-10 (Salmon, Moraes, Dror,
Shaw, SC'11) with counter (grain, step lo, step hi, 0) and key (seed lo, seed hi), then
Box-Muller on the first two word pairs.  Pinned by the published known-answer vectors in
test_philox.py; used by the GPU tests to check the device generator."""
import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(x, dtype=np.uint64) & MASK for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0 & MASK), np.uint64(k1 & MASK)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = np.uint64((int(k0) + W0) & MASK)
        k1 = np.uint64((int(k1) + W1) & MASK)
    return [x.astype(np.uint32) for x in c]


def philox_normals(seed, step0, n_steps, G):
    out = np.zeros((n_steps, G, 3))
    g = np.arange(G, dtype=np.uint64)
    for it in range(n_steps):
        st = step0 + it
        c = philox4x32_10(g, np.full(G, st & MASK), np.full(G, st >> 32), np.zeros(G), seed & MASK, seed >> 32)
        inv = 2.0 ** -32
        u1 = (c[0].astype(np.float64) + 1.0) * inv
        a1 = c[1].astype(np.float64) * inv * 6.283185307179586
        u2 = (c[2].astype(np.float64) + 1.0) * inv
        a2 = c[3].astype(np.float64) * inv * 6.283185307179586
        r1, r2 = np.sqrt(-2.0 * np.log(u1)), np.sqrt(-2.0 * np.log(u2))
        out[it, :, 0] = r1 * np.cos(a1)
        out[it, :, 1] = r1 * np.sin(a1)
        out[it, :, 2] = r2 * np.cos(a2)
    return out
