"""Symbolic analysis of the iteration matrix for networks solved by the
workgroup-per-system kernel (more than 32 equations).

The Newton matrix of calculate_abundances' CVODE solve
(calculate_abundances.c:207) is M = I - gamma*J with J = df/dx, where f
(odes_*.c:26-148) depends on x both directly and through the electron
abundance x_e = sum_j q_j x_j (odes_*.c:44).  The chain term couples every
row that has an electron reaction to every ion column, which fills a large
dense block.  This module works instead with the bordered matrix of size
neq + 1 whose last unknown is x_e:

    [ I - gamma*J0   -gamma*g ] [ dx  ]   [ r ]
    [   -q^T             1    ] [ dxe ] = [ 0 ]

(J0 = df/dx at fixed x_e, g = df/dx_e, q = charges).  Eliminating dxe gives
back (I - gamma*(J0 + g q^T)) dx = r, so both have the same solution, but
the bordered matrix stays as sparse as the reaction network.

The factorisation is a hybrid fixed by the network alone:
  * a static elimination order (Markowitz, ties to the lowest index) chosen
    once per network on the structural pattern;
  * the first ``nhead`` pivots of that order are eliminated without pivoting
    (the head), leaving a Schur complement on the last ``ntail`` unknowns
    (``ntail`` <= 64: one wavefront holds it, one row per lane);
  * the Schur complement is inverted explicitly by Gauss-Jordan with partial
    pivoting (the tail).

Everything here is structural; the numeric work runs on the device from the
tables that gpugen emits.
"""
from __future__ import annotations

import dataclasses
from typing import List

import numpy as np

from .netgen import Network

TAIL_MAX = 64


def bordered_pattern(net: Network) -> np.ndarray:
    """Structural nonzeros of the bordered matrix, (neq+1) x (neq+1)."""
    n = net.neq
    XE = n
    A = np.zeros((n + 1, n + 1), dtype=bool)
    for r in net.reactions:
        if not r.stoich:
            continue
        for f in (r.fa, r.fb):
            if f <= XE:  # a species or the electron (column XE)
                for row in r.stoich:
                    A[row, f] = True
    q = np.asarray(net.charges())
    A[XE, :n] = q != 0
    np.fill_diagonal(A, True)
    return A


def markowitz_order(A: np.ndarray) -> List[int]:
    """Diagonal-pivot Markowitz order on the unsymmetric pattern: at each
    step eliminate the remaining index with the smallest (row count - 1) *
    (column count - 1), ties to the lowest index, and add its fill."""
    B = A.copy()
    n = B.shape[0]
    alive = np.ones(n, dtype=bool)
    order = []
    for _ in range(n):
        idx = np.nonzero(alive)[0]
        sub = B[np.ix_(idx, idx)]
        rc = sub.sum(axis=1) - 1
        cc = sub.sum(axis=0) - 1
        cost = rc.astype(np.int64) * cc.astype(np.int64)
        v = int(idx[int(np.argmin(cost))])
        rows = idx[B[idx, v]]
        cols = idx[B[v, idx]]
        B[np.ix_(rows, cols)] = True
        alive[v] = False
        order.append(v)
    return order


@dataclasses.dataclass
class Analysis:
    nb: int                 # bordered size neq + 1
    order: List[int]        # elimination order (indices of the bordered matrix)
    nhead: int
    ntail: int
    filled: np.ndarray      # pattern after head elimination, in ORIGINAL indices
    head_level: np.ndarray  # [nhead] elimination level of each head pivot
    fwd_level: np.ndarray   # [nhead] forward-solve level of each head row
    bwd_level: np.ndarray   # [nhead] backward-solve level of each head row

    @property
    def tail(self) -> List[int]:
        return self.order[self.nhead:]

    @property
    def head(self) -> List[int]:
        return self.order[: self.nhead]


def analyse(net: Network, tail_max: int = TAIL_MAX) -> Analysis:
    A = bordered_pattern(net)
    nb = A.shape[0]
    order = markowitz_order(A)
    ntail = min(tail_max, nb)
    nhead = nb - ntail
    pos = np.empty(nb, dtype=np.int64)
    pos[order] = np.arange(nb)
    # symbolic head elimination in the original indexing
    F = A.copy()
    for k in order[:nhead]:
        later = pos > pos[k]
        rows = np.nonzero(F[:, k] & later)[0]
        cols = np.nonzero(F[k, :] & later)[0]
        if len(rows) and len(cols):
            F[np.ix_(rows, cols)] = True
    # levels among head pivots: pivot k waits for every earlier pivot j whose
    # elimination touches row k or column k (L or U link)
    head = order[:nhead]
    hl = np.zeros(nhead, dtype=np.int64)
    fl = np.zeros(nhead, dtype=np.int64)
    for a, k in enumerate(head):
        deps = [b for b in range(a) if F[k, head[b]] or F[head[b], k]]
        hl[a] = (max(hl[b] for b in deps) + 1) if deps else 0
        ldeps = [b for b in range(a) if F[k, head[b]]]
        fl[a] = (max(fl[b] for b in ldeps) + 1) if ldeps else 0
    bl = np.zeros(nhead, dtype=np.int64)
    for a in range(nhead - 1, -1, -1):
        k = head[a]
        udeps = [b for b in range(a + 1, nhead) if F[k, head[b]]]
        bl[a] = (max(bl[b] for b in udeps) + 1) if udeps else 0
    return Analysis(nb, order, nhead, ntail, F, hl, fl, bl)
