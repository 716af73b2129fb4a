"""CPU check of the large-instance kernel's data layout (no GPU needed).

The kernel (anechoic_amd/csrc/anc_large.hpp) never forms J D in the slot order; it multiplies MFMA fragments of a
permuted, fragment-major J' with flip fragments that every workgroup scatters into a shared exchange buffer. This test
replays exactly that addressing in numpy -- the permuted index x' -> slot x, the fragment-major J' written by
permute_large, the D' dwords stored by the producers (row tile, wave row/replica half, lane, M/N tile, byte), the
consumers' B-fragment reads and the MFMA operand layout of v_mfma_i32_16x16x64_i8 -- and checks the result against
J D computed directly in the slot order. It also checks that the (row tile, wave, lane, row) ownership covers every
(slot, replica) exactly once and that each lane's rows are complete RNG streams drawn in round order.
"""
import numpy as np
import pytest


def large_slot(xp, kr):
    f = 64 // kr
    c, rho = xp >> 8, xp & 255
    mt, g, e = rho >> 4, (rho >> 2) & 3, rho & 3
    j = 4 * mt + e
    lane = 4 * f * c + f * g + j // kr
    return 256 * (j % kr) + lane


@pytest.mark.parametrize("kr", [16, 32, 64])
def test_slot_permutation_is_a_bijection(kr):
    npk = 256 * kr
    xs = large_slot(np.arange(npk), kr)
    assert np.array_equal(np.sort(xs), np.arange(npk))


@pytest.mark.parametrize("kr", [16])
def test_fragment_gemm_equals_slot_order_product(kr):
    rng = np.random.default_rng(kr)
    npk = 256 * kr
    n = npk - 37                              # padding slots at the end
    NK, NRT = npk // 64, kr
    J = rng.integers(-2, 3, size=(n, n)).astype(np.int64)
    J = np.triu(J, 1)
    J = J + J.T
    R = 128                                   # one work item: 2 replica halves x 64
    D = rng.choice([-2, 0, 2], size=(n, R)).astype(np.int64)     # flips in slot order
    # ---- permute_large: Jf[c][kk][mt][lane][b] = J[slot(256 c + 16 mt + lane % 16)][slot(64 kk + 16 (lane // 16) + b)]
    sl = large_slot(np.arange(npk), kr)
    Jpad = np.zeros((npk, npk), np.int64)
    Jpad[:n, :n] = J
    c, kk, mt, lane, b = np.meshgrid(np.arange(NRT), np.arange(NK), np.arange(16), np.arange(64), np.arange(16), indexing="ij")
    yp = 256 * c + 16 * mt + (lane & 15)
    xp = 64 * kk + 16 * (lane >> 4) + b
    Jf = Jpad[sl[yp], sl[xp]]
    # ---- producers: workgroup row tile crow, wave (wr, wc), lane (g, ml), M-tile mt (0..7), N-tile nt, byte e
    Dpad = np.zeros((npk, R), np.int64)
    Dpad[:n] = D
    xd = np.full((2, NK, 4, 64, 16), 99, np.int64)   # [wc][kk][nt][lane][byte]
    crow, wr, wc, g, ml, mt8, nt, e = np.meshgrid(np.arange(NRT), np.arange(2), np.arange(2), np.arange(4), np.arange(16),
                                                  np.arange(8), np.arange(4), np.arange(4), indexing="ij")
    rho = 128 * wr + 16 * mt8 + 4 * g + e
    x_perm = 256 * crow + rho
    rep = 64 * wc + 16 * nt + ml
    kkp = 4 * crow + 2 * wr + mt8 // 4
    lane_c = 16 * (mt8 % 4) + ml
    byte = 4 * g + e
    xd[wc, kkp, nt, lane_c, byte] = Dpad[sl[x_perm], rep]
    assert not (xd == 99).any()                      # every byte of the exchange is written exactly by someone
    # ---- consumers: wave (wr, wc) of row tile crow, MFMA C[i][j] += sum_k A[i][k] B[k][j]
    # A[i][k] = lane (k // 16, i), byte k % 16 of Jf[crow][kk][8 wr + mt]; B[k][j] = lane (k // 16, j), byte k % 16
    A = Jf.reshape(NRT, NK, 16, 4, 16, 16)            # [c][kk][mt][g][ml][b]  (lane = 16 g + ml)
    A = A.transpose(0, 1, 2, 4, 3, 5).reshape(NRT, NK, 16, 16, 64)   # [c][kk][mt][i][k]
    Bm = xd.reshape(2, NK, 4, 4, 16, 16).transpose(0, 1, 2, 3, 5, 4).reshape(2, NK, 4, 64, 16)   # [wc][kk][nt][k][j]
    # [c][mt16][wc][nt][i][j], summed over kk (q) and k (float64 products are exact for these small integers)
    C = np.rint(np.einsum("cqmik,wqnkj->cmwnij", A.astype(np.float64), Bm.astype(np.float64), optimize=True)).astype(np.int64)
    # output element (mt, e) of lane (g, ml) = C[4 g + e][ml]: row rho = 16 mt16 + i with i = 4 g + e
    H = np.zeros((npk, R), np.int64)
    cc, mm, ww, nn, ii, jj = np.meshgrid(np.arange(NRT), np.arange(16), np.arange(2), np.arange(4), np.arange(16),
                                         np.arange(16), indexing="ij")
    H[sl[256 * cc + 16 * mm + ii], 64 * ww + 16 * nn + jj] = C
    assert np.array_equal(H[:n], J @ D)


@pytest.mark.parametrize("kr", [16, 32, 64])
def test_lane_rows_are_complete_streams_in_round_order(kr):
    """Each lane's 32 rows per wave row-half: for kr < 64 complete streams (all kr rounds of its RNG lanes) in round
    order; for kr = 64 the two row halves split one stream (each wave draws all 64 rounds and keeps its half)."""
    f = 64 // kr
    seen = {}
    for crow in range(kr):
        for wr in range(2):
            for g in range(4):
                lane0 = 4 * f * crow + f * g
                sl0 = 0 if kr == 64 else (32 * wr) // kr
                order = []
                for jj in range(32):
                    j = 32 * wr + jj
                    sub, k = j // kr, j % kr
                    order.append((lane0 + sub, k))
                    x = 256 * k + lane0 + sub
                    assert large_slot(256 * crow + 128 * wr + 16 * (jj // 4) + 4 * g + (jj % 4), kr) == x
                    seen[x] = seen.get(x, 0) + 1
                if kr < 64:
                    # stream index jj // kr is sub-lane sl0 + jj // kr; rounds 0 .. kr-1 in order
                    for s_ in range(32 // kr):
                        part = order[s_ * kr:(s_ + 1) * kr]
                        assert [p[0] for p in part] == [lane0 + sl0 + s_] * kr
                        assert [p[1] for p in part] == list(range(kr))
                else:
                    assert [p[1] for p in order] == list(range(32 * wr, 32 * wr + 32))
    assert sorted(seen) == list(range(256 * kr)) and set(seen.values()) == {1}
