"""The register sorting network of the wave sorts (wb_wave.h), run on the
CPU with each cross-lane move replaced by the lane permutation its DPP
control or shuffle applies: every stage count P and slot count K the
kernels use must leave the first P keys ascending, carry each payload with
its key, and leave the other slots untouched."""
import numpy as np
import pytest

W = 64


def _quad(sel):
    return np.array([(i & ~3) | sel[i & 3] for i in range(W)])


DPP = {
    0xb1: _quad([1, 0, 3, 2]),
    0x1b: _quad([3, 2, 1, 0]),
    0x4e: _quad([2, 3, 0, 1]),
    0x141: np.array([(i & ~7) | (7 - (i & 7)) for i in range(W)]),
    0x140: np.array([(i & ~15) | (15 - (i & 15)) for i in range(W)]),
    0x128: np.array([(i & ~15) | ((i - 8) & 15) for i in range(W)]),  # row_ror 8
}


def _mv(v, ctrl):
    return v[DPP[ctrl]]


def _shfl_xor(v, j):
    return v[np.arange(W) ^ j]


def lane_mirror(v, g):
    return {2: lambda: _mv(v, 0xb1), 4: lambda: _mv(v, 0x1b), 8: lambda: _mv(v, 0x141),
            16: lambda: _mv(v, 0x140)}.get(g, lambda: _shfl_xor(v, g - 1))()


def lane_xor(v, j):
    return {1: lambda: _mv(v, 0xb1), 2: lambda: _mv(v, 0x4e),
            4: lambda: _mv(_mv(v, 0x141), 0x1b), 8: lambda: _mv(v, 0x128)}.get(
                j, lambda: _shfl_xor(v, j))()


def _cx_side(k, v, ok, ov, lower):
    take = np.where(lower, ok < k, ok > k)
    return np.where(take, ok, k), np.where(take, ov, v)


def _cx_slots(ka, va, kb, vb):
    sw = kb < ka
    return (np.where(sw, kb, ka), np.where(sw, vb, va), np.where(sw, ka, kb),
            np.where(sw, va, vb))


def wave_sort(c, v, P):
    """c, v: K x 64 arrays (slot, lane)."""
    K = c.shape[0]
    live = (P + 63) // 64
    lane = np.arange(W)

    def half_clean(J):
        while J >= 1:
            if J >= 64:
                for s in range(K):
                    s2 = s ^ (J // 64)
                    if s2 > s and s < live:
                        c[s], v[s], c[s2], v[s2] = _cx_slots(c[s], v[s], c[s2], v[s2])
            else:
                lower = (lane & J) == 0
                for s in range(min(K, live)):
                    c[s], v[s] = _cx_side(c[s], v[s], lane_xor(c[s], J), lane_xor(v[s], J), lower)
            J //= 2

    G = 2
    while G <= 64 * K and G <= P:
        if G <= 64:
            lower = (lane & (G // 2)) == 0
            for s in range(min(K, live)):
                c[s], v[s] = _cx_side(c[s], v[s], lane_mirror(c[s], G), lane_mirror(v[s], G),
                                      lower)
        else:
            for s in range(K):
                s2 = s ^ (G // 64 - 1)
                if s2 <= s or s >= live:
                    continue
                ok, ov = lane_mirror(c[s2], 64), lane_mirror(v[s2], 64)
                c[s], v[s], ok, ov = _cx_slots(c[s], v[s], ok, ov)
                c[s2], v[s2] = lane_mirror(ok, 64), lane_mirror(ov, 64)
        if G >= 4:
            half_clean(G // 4)
        G *= 2


def _pow2_ceil(m):
    return 1 if m <= 1 else 1 << (m - 1).bit_length()


def test_lane_maps_are_the_intended_partners():
    i = np.arange(W)
    for g in (2, 4, 8, 16, 32, 64):
        assert np.array_equal(lane_mirror(i, g), i ^ (g - 1))
    for j in (1, 2, 4, 8, 16, 32):
        assert np.array_equal(lane_xor(i, j), i ^ j)


@pytest.mark.parametrize("K", [1, 2, 4, 8, 16])
def test_sorts_first_p_keys(K):
    rng = np.random.default_rng(K)
    sizes = sorted({0, 1, 2, 3, 5, 8, 17, 33, 63, 64, 65, 100, 128, 129, 200, 256, 257,
                    500, 512, 513, 700, 1000, 1024} | {64 * K, 32 * K + 1})
    for m in sizes:
        if m > 64 * K:
            continue
        for trial in range(3):
            keys = np.full(64 * K, 0xFFFFFFFF, dtype=np.uint64)
            keys[:m] = rng.choice(1 << 31, size=m, replace=False)
            c = keys.reshape(K, W).copy()
            v = np.arange(64 * K, dtype=np.uint64).reshape(K, W)
            P = _pow2_ceil(m)
            wave_sort(c, v, P)
            flat_c, flat_v = c.reshape(-1), v.reshape(-1)
            assert np.array_equal(flat_c[:m], np.sort(keys[:m])), (K, m)
            assert np.array_equal(keys[flat_v[:m]], flat_c[:m]), (K, m)
            assert sorted(flat_v.tolist()) == list(range(64 * K))
            live = (P + 63) // 64
            assert np.array_equal(flat_c[64 * live:], keys[64 * live:])
