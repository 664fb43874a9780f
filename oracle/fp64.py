"""FP64 finite-horizon expectations for the walk programs -- TEST INFRASTRUCTURE.

For a walk that starts at hop 0, the expected number of visits (updateInfo
calls, randomwalk.hpp:32) to each vertex is sum_{k<H} x_k with x_0 the start
distribution and x_{k+1} = T(x_k), where T is the same chain the step
programs run:

* follow an edge with probability (1-p) when the vertex has out-edges; the
  neighbour is uniform over the vertex's CSR slots, so duplicate edges count
  with multiplicity (randomwalkwithstop.hpp:35-38);
* otherwise (coin, or a dead end):
  JUMP teleports uniformly (randomwalkwithjump.hpp:49-51), STOP/PROB die
  (randomwalkwithstop.hpp:39-41), RESTART/RESTART_JOINT return to the source
  (randomwalkwithrestartwithjoint.hpp:47-49).

H is L for the capped programs and L-1 per period for RESTART_JOINT (the walk
records hop offsets 0..L-2 and stops when hop % L == L-1,
randomwalkwithrestartwithjoint.hpp:52-54).

first_meeting() gives SimRank's per-offset first-meeting probability of two
independent RESTART_JOINT walks (simrank.cpp:71-84) from the product chain
over vertex pairs.  Pins: closed forms on a two-leaves-and-a-hub graph, and
the bit-exact replay's first meetings within the Monte-Carlo bound, tight
enough that a chain without the restart coin fails it.

expected_visits_sources() is expected_visits for many sources in one dense
propagation; pinned by equality with expected_visits source by source.

return_at() gives graphlet's closed-walk probability per source; pinned by a
directed triangle (q^3 at L = 4, 0 at L = 3), brute-force path enumeration
on a small graph, and the bit-exact replay's counts within the binomial bound.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import JUMP, PROB, RESTART, RESTART_JOINT, STOP


def transition(off: np.ndarray, col: np.ndarray) -> tuple[sp.csr_matrix, np.ndarray]:
    """Row-stochastic transition over the CSR (rows with no edges are zero)
    and the dangling indicator."""
    off = np.asarray(off, dtype=np.int64)
    n = off.shape[0] - 1
    deg = np.diff(off)
    rows = np.repeat(np.arange(n), deg)
    w = np.repeat(np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0), deg)
    P = sp.csr_matrix((w, (rows, np.asarray(col, dtype=np.int64))), shape=(n, n))
    P.sum_duplicates()
    return P, deg == 0


def expected_visits(kind: int, off, col, source: int, walks: float, L: int, p: float) -> np.ndarray:
    """Expected per-vertex visit counts of ``walks`` walks started at ``source``
    with hop 0."""
    P, dangling = transition(off, col)
    PT = P.T.tocsr()
    n = P.shape[0]
    x = np.zeros(n)
    x[source] = float(walks)
    H = L if kind != RESTART_JOINT else max(L - 1, 1)
    if kind == PROB:
        H = 10_000  # uncapped: iterate until the surviving mass is negligible
    acc = np.zeros(n)
    for _ in range(H):
        acc += x
        moving = x * (~dangling)
        nxt = (1.0 - p) * (PT @ moving)
        lost = x.sum() - nxt.sum()  # mass that took the else-branch
        if kind == JUMP:
            nxt += lost / n
        elif kind in (RESTART, RESTART_JOINT):
            nxt[source] += lost
        elif kind in (STOP, PROB):
            pass
        x = nxt
        if kind == PROB and x.sum() < 1e-12 * walks:
            break
    return acc


def expected_visits_sources(kind: int, off, col, sources, walks: float, L: int, p: float) -> np.ndarray:
    """expected_visits for many sources at once: row k holds the expected
    visits of ``walks`` walks from sources[k] (the same chain, the columns of
    one dense n x S state propagated together)."""
    P, dangling = transition(off, col)
    PT = P.T.tocsr()
    n = P.shape[0]
    src = np.asarray(sources, dtype=np.int64)
    cols = np.arange(src.shape[0])
    X = np.zeros((n, src.shape[0]))
    X[src, cols] = float(walks)
    H = L if kind != RESTART_JOINT else max(L - 1, 1)
    if kind == PROB:
        H = 10_000
    keep = (~dangling).astype(np.float64)[:, None]
    acc = np.zeros_like(X)
    for _ in range(H):
        acc += X
        nxt = (1.0 - p) * (PT @ (X * keep))
        lost = X.sum(axis=0) - nxt.sum(axis=0)
        if kind == JUMP:
            nxt += lost[None, :] / n
        elif kind in (RESTART, RESTART_JOINT):
            np.add.at(nxt, (src, cols), lost)
        X = nxt
        if kind == PROB and X.sum() < 1e-12 * walks * src.shape[0]:
            break
    return acc.T.copy()


def expected_visits_start(kind: int, off, col, start: np.ndarray, L: int, p: float) -> np.ndarray:
    """Same as expected_visits for walks whose source equals their start
    vertex, given a start histogram (e.g. one walk per vertex)."""
    P, dangling = transition(off, col)
    PT = P.T.tocsr()
    n = P.shape[0]
    if kind in (RESTART, RESTART_JOINT):
        raise ValueError("restart programs depend on each walk's source; sum per-source results")
    x = np.asarray(start, dtype=np.float64).copy()
    acc = np.zeros(n)
    for _ in range(L):
        acc += x
        moving = x * (~dangling)
        nxt = (1.0 - p) * (PT @ moving)
        if kind == JUMP:
            nxt += (x.sum() - nxt.sum()) / n
        x = nxt
    return acc


def mc_l1_bound(expected: np.ndarray, walks: int, visits_per_walk: float, H: int) -> float:
    """Upper bound on E||p_hat - p||_1 for the normalised visit vector.

    Each walk's visit count at v lies in [0, H], so Var <= E[c^2] <= H * E[c];
    with mu_v the expected visits per walk, E|mean - mu_v| <= sqrt(H mu_v / W)
    and the normalised L1 error is at most sum_v sqrt(H mu_v / W) / visits.
    """
    mu = np.asarray(expected, dtype=np.float64) / float(walks)
    return float(np.sqrt(H * mu / walks).sum() / visits_per_walk)


def l1_normalised(counts: np.ndarray, expected: np.ndarray) -> float:
    c = np.asarray(counts, dtype=np.float64)
    e = np.asarray(expected, dtype=np.float64)
    return float(np.abs(c / c.sum() - e / e.sum()).sum())


def first_meeting(off, col, a: int, b: int, p: float, P: int) -> np.ndarray:
    """Probability that a walk from ``a`` and an independent walk from ``b``,
    both running the restart-with-joint chain over this CSR (follow a uniform
    CSR slot with probability 1-p where the vertex has edges, otherwise return
    to the walk's own start), stand on one vertex at offset t for the first
    time, t = 0..P-1: the expectation of simrank.cpp:71-84's first-meeting
    indicator per walk pair.  Exact product chain over vertex pairs, met pairs
    absorbed (dense n x n: small graphs only)."""
    Pm, _ = transition(off, col)
    T = Pm.toarray()
    n = T.shape[0]

    def kernel(src: int) -> np.ndarray:
        K = (1.0 - p) * T
        K[:, src] += 1.0 - K.sum(axis=1)  # the coin or a dead end: back to the start
        return K

    KA, KB = kernel(a), kernel(b)
    J = np.zeros((n, n))
    J[a, b] = 1.0
    out = np.zeros(P)
    for t in range(P):
        if t:
            J = KA.T @ J @ KB
        out[t] = np.trace(J)
        np.fill_diagonal(J, 0.0)
    return out


def return_at(off, col, L: int, p: float) -> np.ndarray:
    """Per source s, the probability that a STOP walk from s (follow a uniform
    CSR slot with probability 1-p, die at a dead end or on the coin) stands
    on s again at hop L-1: graphlet.cpp:46-52's closed-walk indicator (its
    triangle count for L = 4).  diag(M^(L-1)) with M = (1-p) T."""
    P, _ = transition(off, col)
    M = ((1.0 - p) * P).tocsr()
    n = M.shape[0]
    if L <= 1:
        return np.ones(n)
    A = sp.identity(n, format="csr")
    for _ in range(L - 2):
        A = (A @ M).tocsr()
    return np.asarray(A.multiply(M.T.tocsr()).sum(axis=1)).ravel()
