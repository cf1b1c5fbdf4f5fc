"""Writes the golden fixtures in tests/golden/ with numpy only (no C code,
no reference code): the point-set spec, covariance kernel values, SVD ranks of
the C1 tiles and a small dense Cholesky.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform01(seed, ctr):
    return (splitmix64((seed * 0x632BE59BD9B4E019 + ctr) & M64) >> 11) * 2.0 ** -53


def points(N, ndim, seed=0):
    """Jittered grid of side s = ceil(N^(1/ndim)), Morton order of the integer
    grid coordinates (x in the least significant bit), first N points,
    jitter u ~ U(-0.4, 0.4) grid spacings keyed by (grid index, dim)."""
    s = int(np.ceil(N ** (1.0 / ndim) - 1e-9))
    while s ** ndim < N:
        s += 1
    total = s ** ndim
    coords = []
    rem = np.arange(total, dtype=np.int64)
    for _ in range(ndim):
        coords.append(rem % s)
        rem = rem // s
    code = np.zeros(total, dtype=np.uint64)
    for b in range(21):
        for d in range(ndim):
            code |= ((coords[d].astype(np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * ndim + d)
    order = np.argsort(code, kind="stable")[:N]
    xyz = np.zeros((ndim, N))
    for d in range(ndim):
        for i, g in enumerate(order):
            u = uniform01(seed, int(g) * ndim + d)
            xyz[d, i] = (coords[d][g] + 0.5 + 0.8 * u - 0.4) / s
    return xyz


def kernel(kind, X, Y, sigma2, beta, nugget, diag):
    d2 = ((X[:, :, None] - Y[:, None, :]) ** 2).sum(0)
    if kind == 4:
        K = sigma2 * np.exp(-np.sqrt(d2) / beta)
    else:
        K = sigma2 * np.exp(-d2 / (2 * beta * beta))
    if diag:
        K = K + nugget * np.eye(X.shape[1])
    return K


def dsvfr_rank(S, tol):
    """smallest r with sum_{i>=r} S_i^2 <= tol^2 sum S_i^2"""
    tot = (S ** 2).sum()
    tail = np.cumsum((S ** 2)[::-1])[::-1]
    ok = np.nonzero(tail <= tol * tol * tot)[0]
    return int(ok[0]) if len(ok) else len(S)


def main():
    out = {}
    out["pts2_100"] = points(100, 2, 0)
    out["pts3_64"] = points(64, 3, 0)
    out["pts2_2700"] = points(2700, 2, 0)
    out["pts3_1000_s7"] = points(1000, 3, 7)
    P = out["pts2_2700"]
    out["k_sqexp2_blk"] = kernel(2, P[:, 270:290], P[:, 0:20], 1.0, 0.1, 1e-2, False)
    out["k_sqexp2_diag"] = kernel(2, P[:, 0:20], P[:, 0:20], 1.0, 0.1, 1e-2, True)
    P3 = out["pts3_1000_s7"]
    out["k_exp3_blk"] = kernel(4, P3[:, 100:120], P3[:, 0:20], 1.0, 0.1, 1e-2, False)
    out["k_sqexp3_blk"] = kernel(3, P3[:, 100:120], P3[:, 0:20], 2.0, 0.05, 1e-2, False)
    # SVD ranks of the C1 tiles (relative Frobenius rule and absolute sigma >= 1e-8)
    NB, NT = 270, 10
    rel = np.zeros((NT, NT), dtype=np.int32)
    ab = np.zeros((NT, NT), dtype=np.int32)
    for m in range(NT):
        for n in range(m):
            K = kernel(2, P[:, m * NB:(m + 1) * NB], P[:, n * NB:(n + 1) * NB], 1.0, 0.1, 1e-2, False)
            S = np.linalg.svd(K, compute_uv=False)
            rel[m, n] = dsvfr_rank(S, 1e-8)
            ab[m, n] = int((S >= 1e-8).sum())
    out["c1_rank_rel"] = rel
    out["c1_rank_abs"] = ab
    # dense exact C1: Frobenius norm of the lower triangle (dlansy semantics) and log-det
    A = kernel(2, P, P, 1.0, 0.1, 1e-2, True)
    L = np.linalg.cholesky(A)
    out["c1_logdet"] = np.array([2.0 * np.log(np.diag(L)).sum()])
    out["c1_fro"] = np.array([np.linalg.norm(A)])
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), sorted(out))


if __name__ == "__main__":
    main()
