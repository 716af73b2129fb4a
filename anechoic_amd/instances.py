"""Synthetic Ising instances (vectorised numpy restatements of the reference generators, plus build-defined families).

* ``ising_random`` follows gen_ising random (gpu/multibit_bias_20261008/src/gen_ising.cpp:118-135): splitmix64 stream
  seeded with ``seed ^ 0xA5A5F00D12345678``; for every pair x < y (row-major) two draws r0, r1; J = 0 with probability
  pzero, else |J| = 1 + (r1 >> 1) % M, negative iff r1 & 1 (M = 2^(K-1) - 1); biases from ``seed ^ 0x5B1A5B1A5B1A5B1A``.
  K = 2, pzero = 0 is the dense +-1 complete graph of the N = 2000 and N = 16384 workloads.
* ``erdos_renyi`` and ``torus`` are build-defined (no reference counterpart): int2 couplings and biases uniform on
  {-2, -1, +1, +2}.
* ``uneven_batch`` draws the 32-instance uneven workload (N from {800, 1000, 2000, 3000}, half ER at 0.4-6 % density,
  half 2-D tori).
The splitmix64 stream is a counter-based function, so every draw is computed independently in numpy.
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix_draws(state, first, count):
    """Draws number first .. first+count-1 (0-based) of the splitmix64 stream started at `state`."""
    k = np.arange(first + 1, first + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(np.uint64(state) + k * GAMMA)


def ising_random(seed, n, kbits=2, pzero=0.0, bmax=0, ld=None, chunk=1 << 23):
    """gen_ising random: returns (J (n x ld int8, zero padded), bias (n int32))."""
    ld = ld or n
    M = (1 << (kbits - 1)) - 1
    st = np.uint64(seed) ^ np.uint64(0xA5A5F00D12345678)
    lim = np.ldexp(pzero, 64)
    zthr = np.uint64(0xFFFFFFFFFFFFFFFF) if lim >= 18446744073709551615.0 else np.uint64(int(lim))
    J = np.zeros((n, ld), np.int8)
    # pairs (x, y), x < y, in row-major order: pair p uses draws 2p (r0) and 2p + 1 (r1)
    row_start = np.zeros(n + 1, np.int64)
    row_start[1:] = np.cumsum(np.arange(n - 1, -1, -1, dtype=np.int64))
    total = int(row_start[n])
    x0 = 0
    while x0 < n - 1:
        x1 = x0
        while x1 < n - 1 and row_start[x1 + 1] - row_start[x0] <= chunk:
            x1 += 1
        x1 = max(x1, x0 + 1)
        p0, p1 = int(row_start[x0]), int(row_start[x1])
        d = splitmix_draws(st, 2 * p0, 2 * (p1 - p0))
        r0, r1 = d[0::2], d[1::2]
        mag = (1 + ((r1 >> np.uint64(1)) % np.uint64(M))).astype(np.int16)
        v = np.where((r1 & np.uint64(1)) != 0, -mag, mag)
        if pzero > 0:
            v = np.where(r0 < zthr, 0, v)
        rows = np.repeat(np.arange(x0, x1), np.arange(n - 1 - x0, n - 1 - x1, -1))
        cols = np.arange(p0, p1) - row_start[rows] + rows + 1
        J[rows, cols] = v.astype(np.int8)
        J[cols, rows] = v.astype(np.int8)
        x0 = x1
    assert total == 0 or x0 == n - 1
    sb = np.uint64(seed) ^ np.uint64(0x5B1A5B1A5B1A5B1A)
    if bmax > 0:
        d = splitmix_draws(sb, 0, n)
        b = (d % np.uint64(2 * bmax + 1)).astype(np.int64) - bmax
    else:
        b = np.zeros(n, np.int64)
    return J, b.astype(np.int32)


def ising_random_device(seed, n, kbits=2, pzero=0.0, bmax=0, ld=None, device="cuda", stream=None):
    """gen_ising random generated on the GPU (anc_gen_ising): returns torch tensors (J (n x ld int8), bias (n int32)),
    identical to ising_random."""
    import ctypes
    import torch
    from . import _native as N
    ld = ld or n
    J = torch.empty((n, ld), dtype=torch.int8, device=device)
    b = torch.empty(n, dtype=torch.int32, device=device)
    st = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
    N.check(N.lib().anc_gen_ising(ctypes.c_uint64(int(seed) & (2**64 - 1)), int(n), int(kbits), float(pzero), int(bmax),
                                  ctypes.c_void_p(J.data_ptr()), int(ld), ctypes.c_void_p(b.data_ptr()), st))
    return J, b


def _int2(draws):
    """Uniform on {-2, -1, +1, +2} from the low two bits."""
    t = (draws & np.uint64(3)).astype(np.int8)
    return np.array([-2, -1, 1, 2], np.int8)[t]


def erdos_renyi(seed, n, density, ld=None):
    """Build-defined ER graph: pair (x<y) present iff draw_a < density 2^64; J, b uniform on {-2,-1,+1,+2}."""
    ld = ld or n
    st = np.uint64(seed) ^ np.uint64(0xE2D05E1D0000ABCD)
    iu = np.triu_indices(n, 1)
    npairs = len(iu[0])
    d = splitmix_draws(st, 0, 2 * npairs)
    thr = np.uint64(min(int(density * 2.0**64), 2**64 - 1))
    keep = d[0::2] < thr
    J = np.zeros((n, ld), np.int8)
    v = _int2(d[1::2] >> np.uint64(7))
    J[iu[0][keep], iu[1][keep]] = v[keep]
    J[iu[1][keep], iu[0][keep]] = v[keep]
    b = _int2(splitmix_draws(st ^ np.uint64(0xB1A5), 0, n) >> np.uint64(11)).astype(np.int32)
    return J, b


def torus(seed, Lx, Ly, ld=None):
    """Build-defined 2-D toroidal grid (Lx x Ly spins, degree 4): J, b uniform on {-2,-1,+1,+2}."""
    n = Lx * Ly
    ld = ld or n
    st = np.uint64(seed) ^ np.uint64(0x7025C0DE00001234)
    J = np.zeros((n, ld), np.int8)
    idx = np.arange(n).reshape(Ly, Lx)
    right = np.roll(idx, -1, axis=1)
    down = np.roll(idx, -1, axis=0)
    a = np.concatenate([idx.ravel(), idx.ravel()])
    c = np.concatenate([right.ravel(), down.ravel()])
    v = _int2(splitmix_draws(st, 0, len(a)) >> np.uint64(5))
    J[a, c] = v
    J[c, a] = v
    b = _int2(splitmix_draws(st ^ np.uint64(0xB1A5), 0, n) >> np.uint64(11)).astype(np.int32)
    return J, b


TORUS_SHAPES = {800: (32, 25), 1000: (40, 25), 2000: (50, 40), 3000: (60, 50)}


def uneven_batch(seed=20261203, count=32, sizes=(800, 1000, 2000, 3000)):
    """The uneven-size workload: list of dict(kind, n, J, bias, density?)."""
    d = splitmix_draws(np.uint64(seed), 0, 2 * count)
    out = []
    for i in range(count):
        n = sizes[int(d[2 * i] % np.uint64(len(sizes)))]
        if i % 2 == 0:
            dens = 0.004 + (0.06 - 0.004) * float(d[2 * i + 1] >> np.uint64(11)) * 2.0**-53
            J, b = erdos_renyi(seed + 1000 + i, n, dens)
            out.append(dict(kind="er", n=n, density=dens, J=J, bias=b))
        else:
            Lx, Ly = TORUS_SHAPES[n]
            J, b = torus(seed + 1000 + i, Lx, Ly)
            out.append(dict(kind="torus", n=n, J=J, bias=b))
    return out


# ---------------------------------------------------------------------------------------------------- problem encodings
# Lucas-form encodings of gen_ising gp / tsp (gpu/multibit_bias_20261008/src/gen_ising.cpp:145-283), built by the
# library's C++ host code with the reference tool's random streams; K-bit quantisation as problems.py:159-179.

def graph_partition(seed, n, deg=3, k=None, A=1, B=1):
    """Graph partitioning into parts of k and n - k vertices of a random deg-regular graph: (J (n x n int8),
    bias (n int32), meta) with meta = dict(edges (E x 2, u < v), C, sigma=2): H(s) = 2 f(x) + C,
    f = B cut(x) + A (sum x - k)^2, x = (1 + s)/2."""
    import ctypes
    from . import _native as N
    k = n // 2 if k is None else k
    J = np.empty((n, n), np.int8)
    b = np.empty(n, np.int32)
    edges = np.empty((n * deg // 2, 2), np.int32)
    C = ctypes.c_int64()
    N.check(N.lib().anc_gen_gp(int(seed), n, deg, k, int(A), int(B), J.ctypes.data, n, b.ctypes.data, edges.ctypes.data,
                               ctypes.byref(C)))
    return J, b, dict(edges=edges, C=C.value, sigma=2, k=k, A=A, B=B)


def tsp(seed, cities, wmax=2, A=1, B=1):
    """TSP on cities with random distances 1..wmax, city 0 fixed at position 0: (J, bias, meta) with meta =
    dict(W (cities x cities), C, sigma=4); spin (v-1)(cities-1) + (p-1) = city v at position p; H = 4 f + C."""
    import ctypes
    from . import _native as N
    nv = (cities - 1) ** 2
    J = np.empty((nv, nv), np.int8)
    b = np.empty(nv, np.int32)
    W = np.empty((cities, cities), np.int32)
    C = ctypes.c_int64()
    N.check(N.lib().anc_gen_tsp(int(seed), cities, wmax, int(A), int(B), J.ctypes.data, nv, b.ctypes.data, W.ctypes.data,
                                ctypes.byref(C)))
    return J, b, dict(W=W, C=C.value, sigma=4, A=A, B=B)


def gp_cut(edges, spins_pm):
    """Cut size and part size (number of +1 spins) of +-1 spin rows (R x n) for a partitioning instance."""
    s = np.atleast_2d(spins_pm)
    cut = (s[:, edges[:, 0]] != s[:, edges[:, 1]]).sum(1)
    return cut, (s > 0).sum(1)


def tsp_tour_length(W, spins_pm):
    """Tour length of every +-1 spin row that encodes a permutation (city 0 at position 0), -1 otherwise."""
    W = np.asarray(W)
    m = W.shape[0] - 1
    x = np.atleast_2d(spins_pm).reshape(-1, m, m) > 0          # [replica][city v-1][position p-1]
    ok = np.all(x.sum(1) == 1, axis=1) & np.all(x.sum(2) == 1, axis=1)
    out = np.full(x.shape[0], -1, np.int64)
    for r in np.nonzero(ok)[0]:
        tour = np.argmax(x[r], axis=0) + 1                     # city at each position 1..m
        out[r] = int(W[0, tour[0]] + W[tour[-1], 0] + W[tour[:-1], tour[1:]].sum())
    return out


def quantize(J, b, K):
    """Signed K-bit quantisation (problems.py:159-179): alpha = (2^(K-1) - 1) / max|J|, J' = round(alpha J),
    b' = round(alpha b), rounding half away from zero, entrywise. Returns (J' int8, b' int32, alpha)."""
    J = np.asarray(J, np.float64)
    b = np.asarray(b, np.float64)
    mx = float(np.abs(J).max())
    if not 2 <= K <= 8 or mx == 0:
        raise ValueError("quantize: 2 <= K <= 8 and a nonzero J")
    alpha = (2 ** (K - 1) - 1) / mx

    def rnd(x):
        return np.sign(x) * np.floor(np.abs(x) + 0.5)

    return rnd(alpha * J).astype(np.int8), rnd(alpha * b).astype(np.int32), alpha
