"""HIP table generator for the workgroup-per-system BDF kernel (networks with
more than 32 equations: the full network, 214 equations, and MYNETWORK, 57).

Emits ``csrc/gen/wg_<name>.h``.  One stiff system runs on a workgroup of
BLOCK = 256 lanes (four wavefronts).  The tables replace, for these
networks, the reference's MakeRates-generated odes_*.c / jacobian_*.c
(odes_full.c:26-700) and CVODE's dense LU (calculate_abundances.c:207):

* RHS: flux slot s = round*BLOCK + lane (lane-parallel fluxes); the
  per-species sums ydot_i = sum_s S_is flux_s are cut into chunks, every
  chunk owned by one "virtual lane" (VL virtual lanes per lane), so that a
  heavy row (H2, H) is spread over many lanes; a chunk's partial goes to
  P[v] and row i adds its contiguous run of partials.
* Jacobian: exact, of the bordered system (pdrchem.sparse): entries of
  J0 = df/dx at fixed x_e and of g = df/dx_e, each summed by one lane from
  the flux partials D[2s] = dflux_s/dX[a_s], D[2s+1] = dflux_s/dX[b_s].
* Factorisation (pdrchem.sparse): the iteration matrix lives in LDS as
  Hv[NH] (every stored entry whose row or column is a head pivot) followed by
  the dense tail block Tl[NT][NT]; head pivots are eliminated level by level
  without pivoting (phase A: multipliers and pivot reciprocals, phase B:
  updates grouped by target so that one lane owns every update of an entry
  within a level); the tail block is inverted by Gauss-Jordan in wavefront 0.
* Triangular solves of the head by level (forward over L, backward over U),
  the tail by the explicit inverse.
"""
from __future__ import annotations

import heapq
import math
import os
import sys
from typing import Dict, List, Tuple

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from pdrchem.netgen import load_network, Network  # type: ignore
    from pdrchem.gpugen import emit_rate_params, _fmt_list  # type: ignore
    from pdrchem import sparse  # type: ignore
else:
    from .netgen import load_network, Network
    from .gpugen import emit_rate_params, _fmt_list
    from . import sparse

BLOCK = 256
POS_BITS = 14          # LU positions (NLU + 1 < 16384)
MGROUP = 8             # rounds per load group of the M build
SU = 4                 # rounds per load group of the setup stream
# M-build position flags (bits 16.. of an mpos word; row | col << 8 below):
MF_J = 1 << 16         #   the Jacobian slot holds a value here (J entry / charge)
MF_BORD = 1 << 17      #   the bordered row: coefficient -1 in place of -gamma
MF_DIAG = 1 << 18      #   add 1 (M = I - gamma J)
MF_FIXED = 1 << 19     #   the zero or one slot: unscaled constant
MF_ONE = 1 << 20       #   ... which is 1
MF_SKIP = 1 << 21      #   padding: nothing to store
SCHUR_G32 = 4          # words per load group of the FP32 Schur loop
SCHUR_AHEAD = 2        # load groups in flight ahead of the Schur loops' use
COEF_BITS = 4          # signed stoichiometric coefficients (-8..7)


def _lpt(loads: List[Tuple[int, object]], nbins: int) -> List[List[object]]:
    """Longest-processing-time assignment of (weight, item) to nbins lanes;
    deterministic (ties to the lower lane)."""
    bins: List[List[object]] = [[] for _ in range(nbins)]
    heap = [(0, b) for b in range(nbins)]
    heapq.heapify(heap)
    for w, item in sorted(loads, key=lambda x: (-x[0], str(x[1]))):
        load, b = heapq.heappop(heap)
        bins[b].append(item)
        heapq.heappush(heap, (load + w, b))
    return bins


def stream_word(a: int, b: int, c: int, t: int, last: int, store: int = 0) -> int:
    """A setup-stream word: four LU positions as LDS byte offsets of 4-byte
    entries (position * 4, 16 bits each: a | b << 16 | c << 32 | t << 48),
    `last` in bit 48 and `store` in bit 49 (the target offset's low bits,
    which are zero); bit 0 (the first offset's low bit) is left for the
    phase-end flag.  Elimination: operands a, b, pivot c; inverse: src1 a,
    src2 b, scale c."""
    for x in (a, b, c, t):
        if not 0 <= x < (1 << POS_BITS):
            raise ValueError("LU position out of range")
    return (4 * a) | ((4 * b) << 16) | ((4 * c) << 32) | ((4 * t) << 48) | (last << 48) | (store << 49)


def word_fields(w: int):
    """(a, b, c, t, last, store, phase_end) of a setup-stream word."""
    return ((w & 0xFFFC) >> 2, ((w >> 16) & 0xFFFC) >> 2, ((w >> 32) & 0xFFFC) >> 2, ((w >> 48) & 0xFFFC) >> 2,
            (w >> 48) & 1, (w >> 49) & 1, w & 1)


def _check_phase(lanes: List[List[int]], elim: bool) -> None:
    """Within one barrier-separated phase no target may be written twice or
    read as an operand by another item (lanes run concurrently)."""
    tgts, srcs = set(), set()
    for lst in lanes:
        for w in lst:
            a, b, c, t, last, _, _ = word_fields(w)
            # an elimination store rescales its own entry (reads and writes t)
            srcs.update(x for x in (a, b, c) if not (elim and x == t))
            if last:
                if t in tgts:
                    raise ValueError("a phase writes a position twice")
                tgts.add(t)
    if tgts & srcs:
        raise ValueError("a phase reads a position it writes")


def rhs_word32(w16: int) -> int:
    """Kernel word of a packed RHS term (slot | 4-bit coefficient << 12): the
    flux's LDS byte offset and the coefficient as an FP64 high half."""
    from pdrchem.gpugen import rhs_word
    slot, c = w16 & 0xFFF, (w16 >> 12) & 0xF
    return rhs_word(slot, c - 16 if c >= 8 else c)


def _coef4(c: int) -> int:
    if not -8 <= c <= 7:
        raise ValueError(f"stoichiometric coefficient {c} does not fit 4 bits")
    return c & 0xF


class WgTables:
    def __init__(self, net: Network, block: int = BLOCK):
        self.net = net
        self.block = block
        n = net.neq
        self.neq = n
        an = sparse.analyse(net)
        self.an = an
        NB = an.nb
        self.nb = NB
        if NB > 255:
            raise ValueError(f"{net.name}: more than 254 equations")
        XE, ONE = n, n + 1
        # ------------------------------------------------------------ fluxes
        self.slot_reac = [r.index for r in net.reactions if r.stoich]
        nact = len(self.slot_reac)
        self.fr = max(1, math.ceil(nact / block))
        self.nslot = self.fr * block
        reac_slot = {r: s for s, r in enumerate(self.slot_reac)}
        self.slot_fab = []
        for s in range(self.nslot):
            if s < nact:
                r = net.reactions[self.slot_reac[s]]
                self.slot_fab.append(r.fa | (r.fb << 8) | (r.nh << 16) | (1 << 17))
            else:
                self.slot_fab.append(ONE | (ONE << 8))
        self.fzero = self.nslot                 # F[NSLOT] = 0
        # -------------------------------------------------------- RHS chunks
        rows: List[List[Tuple[int, int]]] = [[] for _ in range(n)]
        for r in net.reactions:
            if not r.stoich:
                continue
            s = reac_slot[r.index]
            for k, v in sorted(r.stoich.items()):
                rows[k].append((s, v))
        best = None
        for vl in (1, 2, 3, 4):
            nv = vl * block
            nl = [1 if rows[i] else 0 for i in range(n)]
            if sum(nl) > nv:
                continue
            load = lambda i: math.ceil(len(rows[i]) / nl[i]) if nl[i] else 0
            heap = [(-load(i), i) for i in range(n) if nl[i]]
            heapq.heapify(heap)
            used = sum(nl)
            while used < nv and heap:
                negl, i = heapq.heappop(heap)
                if -negl <= 1:
                    heapq.heappush(heap, (negl, i))
                    break
                nl[i] += 1
                used += 1
                heapq.heappush(heap, (-load(i), i))
            tm = max(load(i) for i in range(n))
            cost = vl * tm + vl * 4  # chunk passes plus per-pass overhead
            if best is None or cost < best[0]:
                best = (cost, vl, nl[:], tm)
        _, self.vl, nl, self.tm = best
        self.nv = self.vl * block
        self.row_nl = nl
        base, b = [], 0
        for i in range(n):
            base.append(b)
            b += nl[i]
        self.row_base = base
        self.pm = max(nl)
        self.pzero = self.nv                   # P[NV] = 0
        zt = self.fzero | (0 << 12)
        self.rhs_term = [[[zt] * block for _ in range(self.tm)] for _ in range(self.vl)]
        for i in range(n):
            terms = rows[i]
            if not nl[i]:
                continue
            csz = math.ceil(len(terms) / nl[i])
            for u in range(nl[i]):
                v = base[i] + u
                pas, lane = divmod(v, block)
                for t, (s, c) in enumerate(terms[u * csz:(u + 1) * csz]):
                    if s >= 4096:
                        raise ValueError("flux slot does not fit 12 bits")
                    self.rhs_term[pas][t][lane] = s | (_coef4(c) << 12)
        # ------------------------------------------------------- Jacobian
        jterms: Dict[Tuple[int, int], List[Tuple[int, int]]] = {}
        for r in net.reactions:
            if not r.stoich:
                continue
            s = reac_slot[r.index]
            for i, coef in r.stoich.items():
                for which, f in ((0, r.fa), (1, r.fb)):
                    if f <= XE:
                        jterms.setdefault((i, f), []).append((2 * s + which, coef))
        self.jent = sorted(jterms)             # (row, col) with col == XE for g
        self.nj = len(self.jent)
        jid = {e: k for k, e in enumerate(self.jent)}
        self.dzero = 2 * self.nslot            # D[2*NSLOT] = 0
        bins = _lpt([(len(jterms[e]), e) for e in self.jent], block)
        lists = []
        for lane in range(block):
            lst = []
            for e in sorted(bins[lane]):
                tl = jterms[e]
                for t, (d, c) in enumerate(tl):
                    last = 1 if t == len(tl) - 1 else 0
                    lst.append(d | (_coef4(c) << 13) | (last << 17) | (jid[e] << 18))
            lists.append(lst)
        self.jtm = max(len(l) for l in lists)
        if self.dzero >= (1 << 13) or self.nj >= (1 << 14):
            raise ValueError("Jacobian tables overflow their bit fields")
        pad = self.dzero
        self.jac_term = [[(lists[lane][t] if t < len(lists[lane]) else pad) for lane in range(block)]
                         for t in range(self.jtm)]
        # --------------------------------------------------- LU positions
        F = an.filled
        order = an.order
        head, tail = an.head, an.tail
        self.nh_piv = an.nhead
        self.nt = an.ntail
        if self.nt > 64:
            raise ValueError("tail wider than a wavefront")
        pos_of = np.full((NB, NB), -1, dtype=np.int64)
        is_head = np.zeros(NB, dtype=bool)
        is_head[head] = True
        hv = [(i, j) for i in range(NB) for j in range(NB) if F[i, j] and (is_head[i] or is_head[j])]
        hv.sort(key=lambda e: (e[0], e[1]))
        for p, (i, j) in enumerate(hv):
            pos_of[i, j] = p
        self.nhv = len(hv)
        self.luzero = self.nhv                 # LU[NHV] = 0
        self.luone = self.nhv + 1              # LU[NHV+1] = 1
        # explicit inverses of the head triangles (solves without levels):
        # Linv = strict lower part of L_HH^-1, Winv = U_HH^-1 (with diagonal)
        h = an.nhead
        hpos = {v: a for a, v in enumerate(head)}
        lrow = [[b for b in range(a) if F[head[a], head[b]]] for a in range(h)]
        urow = [[b for b in range(a + 1, h) if F[head[a], head[b]]] for a in range(h)]
        li_pat = [set() for _ in range(h)]
        for a in range(h):
            for b in lrow[a]:
                li_pat[a].add(b)
                li_pat[a] |= li_pat[b]
        wi_pat = [set() for _ in range(h)]
        for a in range(h - 1, -1, -1):
            wi_pat[a].add(a)
            for b in urow[a]:
                wi_pat[a] |= wi_pat[b]
        self.hi_base = self.nhv + 2
        li_idx, wi_idx = {}, {}
        k = self.hi_base
        for a in range(h):
            for cc in sorted(li_pat[a]):
                li_idx[(a, cc)] = k
                k += 1
        for a in range(h):
            for cc in sorted(wi_pat[a]):
                wi_idx[(a, cc)] = k
                k += 1
        self.nhi = k - self.hi_base
        # tail block Tl[NT][NT] after these, 16-byte aligned: it lies over the
        # flux array F in LDS, which is FP64
        self.tbase = 4 * math.ceil(k / 4)
        # row stride NT + 1: a column of the block falls in distinct LDS banks
        self.tstride = self.nt + 1
        tidx = {v: r for r, v in enumerate(tail)}
        for i in tail:
            for j in tail:
                pos_of[i, j] = self.tbase + tidx[i] * self.tstride + tidx[j]
        self.nlu = self.tbase + self.nt * self.tstride
        if self.nlu >= (1 << POS_BITS):
            raise ValueError("LU storage exceeds the position field")
        # the saved Jacobian is kept by LU position, so that building M reads
        # it in the same order as the position table: the Jacobian terms'
        # entry field becomes the entry's LU position
        self.jpos = [int(pos_of[i, j]) for (i, j) in self.jent]
        if min(self.jpos) < 0:
            raise ValueError("a Jacobian entry has no LU position")
        self.jac_term = [[(w & 0x3FFFF) | (self.jpos[w >> 18] << 18) for w in row] for row in self.jac_term]
        self.tail = tail
        self.head = head
        allpos = hv + [None, "one"] + ["pad"] * (self.tbase - self.nhv - 2)
        for i in tail:
            allpos += [(i, j) for j in tail] + ["pad"]
        assert len(allpos) == self.nlu
        # M entry descriptors in position order: row | col << 8 | flags (see
        # MF_*; the zero and one slots and padding have row = col = 0)
        self.mpos = []
        for e in allpos:
            if e is None:
                self.mpos.append(MF_FIXED)
                continue
            if e == "pad":
                self.mpos.append(MF_SKIP)
                continue
            if e == "one":
                self.mpos.append(MF_FIXED | MF_ONE)
                continue
            i, j = e
            f = 0
            if i < n:
                f |= MF_J if (i, j) in jid else 0
            else:                     # the bordered row: charges, and 1 at x_e
                f |= MF_BORD | (MF_J if j < n else 0)
            f |= MF_DIAG if i == j else 0
            self.mpos.append(i | (j << 8) | f)
        self.mtm = math.ceil(self.nlu / block)
        # padded to whole groups of the M build's loads (MGROUP rounds)
        self.mtmp = MGROUP * math.ceil(self.mtm / MGROUP)
        self.mpos += [MF_SKIP] * (self.mtmp * block - self.nlu)
        # ------------------------------------------ head elimination levels
        hl = an.head_level
        self.nlev = int(hl.max()) + 1 if an.nhead else 0
        later_mask = {}
        rank = {v: a for a, v in enumerate(order)}
        # Elimination phases.  Phase lv (0..NLEV) runs, in one barrier-
        # separated step, the updates of the level-lv pivots and the stores of
        # the level-(lv-1) pivots' multipliers and diagonal reciprocals.  An
        # update term recomputes its multiplier a_ik / a_kk from the unscaled
        # entry (final after phase lv-1; the store of L_ik happens only in
        # phase lv+1), so the two never meet in one phase.  Item word (see
        # stream_word): operands a, b, p and target t; m = a / p (hardware
        # reciprocal in the FP32 form); acc += m * b; at `last`:
        # t <- store ? m : t - acc.
        zero, one = self.luzero, self.luone
        word = lambda a, b, p_, t, last, store: stream_word(a, b, p_, t, last, store)
        stores: List[List[Tuple[int, object, List[int]]]] = []   # per level: (weight, key, words)
        runs: List[List[Tuple[int, object, List[int]]]] = []
        for lev in range(self.nlev):
            piv = [head[a] for a in range(an.nhead) if hl[a] == lev]
            st, upd = [], {}
            for k in piv:
                dk = int(pos_of[k, k])
                st.append((1, ('d', k), [word(one, zero, dk, wi_idx[(hpos[k], hpos[k])], 1, 1)]))
                rows_l = [i for i in range(NB) if F[i, k] and rank[i] > rank[k]]
                cols_u = [j for j in range(NB) if F[k, j] and rank[j] > rank[k]]
                for i in rows_l:
                    lp = int(pos_of[i, k])
                    st.append((1, ('l', i, k), [word(lp, zero, dk, lp, 1, 1)]))
                    for j in cols_u:
                        if not (is_head[i] or is_head[j]):
                            continue  # tail block: Schur complement product
                        tp = int(pos_of[i, j])
                        assert 0 <= tp < self.nhv, (i, j)
                        upd.setdefault(tp, []).append((lp, int(pos_of[k, j]), dk))
            stores.append(st)
            rl = []
            for tp in sorted(upd):
                us = upd[tp]
                rl.append((len(us), ('u', tp), [word(lp, up, dk, tp, 1 if t == len(us) - 1 else 0, 0)
                                                for t, (lp, up, dk) in enumerate(us)]))
            runs.append(rl)
        pad_e = word(zero, zero, one, zero, 0, 0)
        self.elim_phases: List[List[List[int]]] = []   # [phase][lane] -> words
        for ph in range(self.nlev + 1 if self.nlev else 0):
            jobs = (runs[ph] if ph < self.nlev else []) + (stores[ph - 1] if ph > 0 else [])
            bins = _lpt([(w, (key, words)) for (w, key, words) in jobs], block)
            lanes = []
            for lane in range(block):
                lst = []
                for key, words in sorted(bins[lane], key=lambda kw: str(kw[0])):
                    lst += words
                lanes.append(lst)
            _check_phase(lanes, elim=True)
            self.elim_phases.append(lanes)
        self.pad_e = pad_e
        # ------------------------- Schur complement / tail coupling tables
        # KH head pivots padded to a multiple of 4 (MFMA K steps); the
        # multipliers L[t][k] of tail rows and U[k][j] of tail columns
        if an.nhead and self.nt != 64:
            raise ValueError("a head needs a 64-wide tail (one MFMA tile grid)")
        self.kh = 4 * math.ceil(an.nhead / 4)
        self.head_row = list(head) + [NB] * (self.kh - an.nhead)   # B[NB] = 0
        self.lth = [self.luzero] * (64 * self.kh)                 # [t][k]
        self.uht = [self.luzero] * (self.kh * 64)                 # [k][j]
        for a, k in enumerate(head):
            for t, i in enumerate(tail):
                if F[i, k]:
                    self.lth[t * self.kh + a] = int(pos_of[i, k])
                if F[k, i]:
                    self.uht[a * 64 + t] = int(pos_of[k, i])
        self.schur32, self.schur64 = self._schur_tables()
        # ------------------------ explicit head inverses, computed in the setup
        # Linv[a][c] = -(L[a][c] + sum_b L[a][b] Linv[b][c])  (rows by forward level)
        # Winv[a][c] = -Winv[a][a] * sum_b U[a][b] Winv[b][c]  (rows by backward level)
        # term: src1 | src2 << 14 | target << 28 | scale << 42 | last << 56;
        # target <- -(sum of src1*src2) * LU[scale]
        fl, bl = an.fwd_level, an.bwd_level
        self.inv_levels: List[List[List[int]]] = []
        nfl = int(fl.max()) + 1 if h else 0
        nbl = int(bl.max()) + 1 if h else 0
        lpos = lambda a, b: int(pos_of[head[a], head[b]])
        for lev in range(nfl):
            per_t = {}
            for a in range(h):
                if fl[a] != lev:
                    continue
                for cc in sorted(li_pat[a]):
                    terms = []
                    if cc in lrow[a]:
                        terms.append((lpos(a, cc), self.luone))
                    for b in lrow[a]:
                        if cc in li_pat[b]:
                            terms.append((lpos(a, b), li_idx[(b, cc)]))
                    per_t[(li_idx[(a, cc)], self.luone)] = terms
            self.inv_levels.append(self._term_tab(per_t))
        for lev in range(nbl):
            per_t = {}
            for a in range(h):
                if bl[a] != lev:
                    continue
                for cc in sorted(wi_pat[a]):
                    if cc == a:
                        continue
                    terms = [(lpos(a, b), wi_idx[(b, cc)]) for b in urow[a] if cc in wi_pat[b]]
                    per_t[(wi_idx[(a, cc)], wi_idx[(a, a)])] = terms
            self.inv_levels.append(self._term_tab(per_t))
        # ------------------------------- solve: four chunked sparse mat-vecs
        # S1 y_H = b_H + Linv b_H, S2 tail gather L_TH y_H, S4 U_HT x_T,
        # S5 x_H = Winv z_H; entry: LU position | B index << 14
        s1 = [[(li_idx[(a, cc)], head[cc]) for cc in sorted(li_pat[a])] for a in range(h)]
        s2 = [[(int(pos_of[i, head[b]]), head[b]) for b in range(h) if F[i, head[b]]] for i in tail]
        s4 = [[(int(pos_of[head[a], j]), j) for j in tail if F[head[a], j]] for a in range(h)]
        s5 = [[(wi_idx[(a, cc)], head[cc]) for cc in sorted(wi_pat[a])] for a in range(h)]
        self.spmv = [self._spmv_tab(rows) for rows in (s1, s2, s4, s5)]
        self.n_inv_terms = sum(sum(1 for row in lev for v in row if word_fields(v)[4]) for lev in self.inv_levels)
        # the head pivots' final diagonal entries and stored reciprocals,
        # checked after the elimination: dpos * 4 | rpos * 4 << 16
        self.piv_chk = [4 * int(pos_of[k, k]) | (4 * wi_idx[(hpos[k], hpos[k])]) << 16 for k in head]
        self._build_stream()
        self.diag_pos = [int(pos_of[i, i]) for i in range(NB)]
        self.charges = net.charges()

    def _build_stream(self) -> None:
        """The setup's elimination and inverse phases as one [round][lane]
        stream of 64-bit words, every phase padded to whole groups of SU
        rounds (so the kernel's loads are unconditional and the group after a
        phase's last is the next phase's first), plus two groups of padding at
        the end for the kernel's prefetch two groups ahead.  Empty phases are
        dropped."""
        block = self.block
        pad_i = stream_word(self.luzero, self.luzero, self.luzero, self.luzero, 0)
        phases = []
        for lanes in self.elim_phases:
            r = max((len(l) for l in lanes), default=0)
            if r:
                r = SU * math.ceil(r / SU)
                phases.append([[lanes[l][t] if t < len(lanes[l]) else self.pad_e for l in range(block)]
                               for t in range(r)])
        self.nep = len(phases)
        for lev in self.inv_levels:
            if not lev:
                continue
            _check_phase([[lev[t][l] for t in range(len(lev))] for l in range(block)], elim=False)
            r = SU * math.ceil(len(lev) / SU)
            phases.append([list(lev[t]) if t < len(lev) else [pad_i] * block for t in range(r)])
        # bit 0 of every lane's word in the first round of a phase's last
        # group: the phase ends after this group (read by the kernel from the
        # words it already holds, instead of a per-phase length)
        for ph in phases:
            ph[len(ph) - SU] = [w | 1 for w in ph[len(ph) - SU]]
        self.SU = SU
        self.sp_groups = [len(ph) // SU for ph in phases]
        self.stream = [w for ph in phases for row in ph for w in row] + [self.pad_e] * (2 * SU * block)

    def _term_tab(self, per_t):
        """Targets with their (src1, src2) terms -> [round][lane] 64-bit items,
        every target's terms in one lane (LPT over lanes)."""
        bins = _lpt([(max(1, len(v)), t) for t, v in per_t.items()], self.block)
        lanes = []
        for lane in range(self.block):
            lst = []
            for key in sorted(bins[lane]):
                tgt, scale = key
                terms = per_t[key] or [(self.luzero, self.luzero)]
                for t, (a, b) in enumerate(terms):
                    last = 1 if t == len(terms) - 1 else 0
                    lst.append(stream_word(a, b, scale, tgt, last))
            lanes.append(lst)
        r = max((len(l) for l in lanes), default=0)
        pad = stream_word(self.luzero, self.luzero, self.luzero, self.luzero, 0)
        return [[(lanes[lane][t] if t < len(lanes[lane]) else pad) for lane in range(self.block)]
                for t in range(r)]

    def _spmv_tab(self, rows):
        """Rows of (LU position, B index) entries -> chunks, one per lane:
        every row gets at least one chunk (also when empty), extra chunks go
        to the longest rows.  Returns dict(tab=[round][lane] words,
        base=[row], cnt=[row]) with partials P[lane] of chunk = lane."""
        nr = len(rows)
        if nr == 0:
            return dict(tab=[], base=[], cnt=[], rounds=0)
        if nr > self.block:
            raise ValueError("more rows than lanes")
        nl = [1] * nr
        heap = [(-len(rows[i]), i) for i in range(nr)]
        heapq.heapify(heap)
        used = nr
        while used < self.block and heap:
            negl, i = heapq.heappop(heap)
            if math.ceil(len(rows[i]) / nl[i]) <= 1:
                break
            nl[i] += 1
            used += 1
            heapq.heappush(heap, (-math.ceil(len(rows[i]) / nl[i]), i))
        base, b = [], 0
        for i in range(nr):
            base.append(b)
            b += nl[i]
        rounds = max(math.ceil(len(rows[i]) / nl[i]) for i in range(nr))
        pad = self.luzero
        tab = [[pad] * self.block for _ in range(rounds)]
        for i in range(nr):
            csz = math.ceil(len(rows[i]) / nl[i]) if rows[i] else 0
            for u in range(nl[i]):
                for t, (p, col) in enumerate(rows[i][u * csz:(u + 1) * csz]):
                    tab[t][base[i] + u] = p | (col << 14)
        return dict(tab=tab, base=base, cnt=nl, rounds=rounds)

    # ------------------------------------------------------------------ emit
    def emit(self, path: str) -> None:
        net = self.net
        ns = f"pdrwg_{net.name}"
        L = [f"// Generated by pdrchem/wggen.py from networks/{net.name}.json; do not edit.",
             "// Tables of the workgroup-per-system BDF kernel (csrc/bdf_wg_impl.h).",
             "#pragma once", "#include <stdint.h>", f"namespace {ns} {{"]
        c = lambda k, v, cm="": L.append(f"constexpr int {k} = {v};" + (f"  // {cm}" if cm else ""))
        c("NSPEC", net.nspec); c("NREAC", net.nreac); c("NEQ", net.neq)
        c("NB", self.nb, "bordered size: species + x_e")
        c("BLOCK", self.block)
        c("FR", self.fr, "flux rounds per lane"); c("NSLOT", self.nslot)
        c("NACTIVE", len(self.slot_reac))
        c("VL", self.vl, "RHS virtual lanes per lane"); c("NV", self.nv)
        c("TM", self.tm, "RHS terms per chunk"); c("PM", self.pm, "chunks per row (max)")
        c("NJ", self.nj, "Jacobian entries (J0 and g)"); c("JTM", self.jtm, "Jacobian terms per lane")
        c("NHV", self.nhv, "head-related LU entries; LU[NHV] = 0")
        c("TBASE", self.tbase, "LU position of the tail block Tl[NT][NT] (multiple of 4)")
        c("NT", self.nt, "tail size")
        c("TS", self.tstride, "row stride of the tail block")
        c("NLU", self.nlu); c("MTM", self.mtm, "M-build rounds")
        c("MTMP", self.mtmp, "M-build rounds padded to whole load groups (the saved Jacobian's size in rounds)")
        c("MGROUP", MGROUP, "rounds per M-build load group")
        for nm_, v_ in (("MF_J", MF_J), ("MF_BORD", MF_BORD), ("MF_DIAG", MF_DIAG), ("MF_FIXED", MF_FIXED),
                        ("MF_ONE", MF_ONE), ("MF_SKIP", MF_SKIP)):
            L.append(f"constexpr uint32_t {nm_} = 0x{v_:x}u;  // M-build position flag")
        c("KH", self.kh, "head pivots padded to a multiple of 4")
        c("SCHUR_G32", SCHUR_G32, "words per load group of the FP32 Schur loop")
        c("SCHUR_AHEAD", SCHUR_AHEAD, "Schur load groups in flight ahead of use")
        c("NHEAD", self.nh_piv); c("NLEV", self.nlev, "head elimination levels")
        c("LUONE", self.luone, "LU[LUONE] = 1")
        c("NHI", self.nhi, "entries of the head inverses Linv, Winv")
        c("NINV", len(self.inv_levels), "levels of the head-inverse computation")
        c("XLAMBDA_SPLINE", net.xlambda_spline)

        def arr(ctype, name, vals, fmt=str, per=16):
            L.append(f"__device__ const {ctype} {name}[{max(1, len(vals))}] = {{")
            L.append(_fmt_list([fmt(v) for v in vals] if vals else ["0"], per))
            L.append("};")

        hx = lambda v: f"0x{v:x}u"
        hx64 = lambda v: f"0x{v:x}ull"
        arr("int16_t", "slot_reac", self.slot_reac + [-1] * (self.nslot - len(self.slot_reac)))
        arr("uint32_t", "slot_fab", self.slot_fab, hx, 8)
        flat = [rhs_word32(self.rhs_term[p][t][l])
                for p in range(self.vl) for t in range(self.tm) for l in range(self.block)]
        L.append("// [pass][t][lane]: byte offset of the flux slot (8 * slot) | top 16 bits of the")
        L.append("// coefficient's FP64 high word << 16 (the rest of the double is zero)")
        arr("uint32_t", "rhs_term", flat, hx, 8)
        rb = [self.row_base[i] if i < self.neq else 0 for i in range(self.block)]
        rn = [self.row_nl[i] if i < self.neq else 0 for i in range(self.block)]
        arr("uint16_t", "row_base", rb)
        arr("uint8_t", "row_nl", rn)
        L.append("// [t][lane]: D index | coef << 13 | last << 17 | LU position of the entry << 18")
        arr("uint32_t", "jac_term", [v for row in self.jac_term for v in row], hx, 8)
        L.append("// entry -> row | col << 8 (col NEQ: d/dx_e)")
        arr("uint16_t", "jent", [i | (j << 8) for (i, j) in self.jent], hx, 12)
        L.append("// [t][lane]: row | col << 8 | J entry << 16 (0xffff: none; the saved Jacobian holds")
        L.append("// the entry at this position); row 255 = padding")
        arr("uint32_t", "mpos", self.mpos, hx, 8)
        L.append("// setup stream [round][lane]: the elimination phases (NEP), then the head-inverse")
        L.append("// phases; every phase a whole number of SU-round groups (sp_groups), one spare group")
        L.append("// at the end.  Word: four LU positions as byte offsets of 4-byte entries (position * 4,")
        L.append("// 16 bits each): a | b << 16 | c << 32 | t << 48; last = bit 48, store = bit 49; bit 0")
        L.append("// (in every lane's word of a group's first round): the phase ends with this group.")
        L.append("// Elimination: m = a / c; acc += m * b; at last: t <- store ? m : t - acc.")
        L.append("// Inverse: acc += a * b; at last: t <- -acc * c.")
        c("SU", SU, "rounds per setup-stream load group")
        c("NEP", self.nep, "elimination phases")
        c("NSP", len(self.sp_groups), "setup-stream phases")
        arr("int", "sp_groups", self.sp_groups)
        arr("uint64_t", "setup_stream", self.stream, hx64, 6)
        L.append("// head pivots: diagonal entry and stored reciprocal, as byte offsets (position * 4):")
        L.append("// dpos | rpos << 16")
        arr("uint32_t", "piv_chk", self.piv_chk, hx, 8)
        for nm, sp in zip(("s1", "s2", "s4", "s5"), self.spmv):
            L.append(f"// solve stage {nm} [round][lane]: LU position | B index << 14")
            c(f"{nm.upper()}_ROUNDS", sp["rounds"])
            c(f"{nm.upper()}_MAXC", max(sp["cnt"], default=0), "most chunks of one row")
            arr("uint32_t", f"{nm}_tab", [v for row in sp["tab"] for v in row], hx, 8)
            arr("uint16_t", f"{nm}_base", sp["base"])
            arr("uint8_t", f"{nm}_cnt", sp["cnt"])
        # per lane of the solve, packed: head row, the four stages' chunk
        # bases (lane tid) and the tail row with its S2 base and every count
        # (tail fields of lane tid & 63)
        def sp_field(k, key, i, default):
            v = self.spmv[k][key]
            return v[i] if i < len(v) else default
        w0, w1 = [], []
        for tid in range(self.block):
            tl = tid & 63
            hr = self.head_row[tid] if tid < len(self.head_row) else self.nb
            trow = self.tail[tl] if tl < self.nt else 0
            b1, b4, b5 = (sp_field(k, "base", tid, 0) for k in (0, 2, 3))
            c1, c4, c5 = (sp_field(k, "cnt", tid, 0) for k in (0, 2, 3))
            b2, c2 = sp_field(1, "base", tl, 0), sp_field(1, "cnt", tl, 0)
            for v, bits in ((hr, 8), (b1, 8), (b4, 8), (b5, 8), (trow, 8), (b2, 8), (c1, 4), (c2, 4), (c4, 4), (c5, 4)):
                if not 0 <= v < (1 << bits):
                    raise ValueError("solve lane field overflows")
            w0.append(hr | (b1 << 8) | (b4 << 16) | (b5 << 24))
            w1.append(trow | (b2 << 8) | (c1 << 16) | (c2 << 20) | (c4 << 24) | (c5 << 28))
        L.append("// solve constants per lane: head_row | s1_base << 8 | s4_base << 16 | s5_base << 24 and")
        L.append("// tail_row | s2_base << 8 | s1_cnt << 16 | s2_cnt << 20 | s4_cnt << 24 | s5_cnt << 28")
        L.append("// (head fields of lane tid, tail fields of lane tid & 63)")
        arr("uint32_t", "solve_lane0", w0, hx, 8)
        arr("uint32_t", "solve_lane1", w1, hx, 8)
        arr("uint8_t", "tail_row", list(self.tail) + [0] * (64 - self.nt))
        arr("uint8_t", "head_row", self.head_row)
        L.append("// LU positions of L[tail t][head k] ([t][k]) and U[head k][tail j] ([k][j])")
        arr("uint16_t", "lth_pos", self.lth, str, 16)
        arr("uint16_t", "uht_pos", self.uht, str, 16)
        L.append("// Schur operand positions per lane, two k steps a word (A | B << 16 | A' << 32 | B' << 48):")
        L.append("// FP32 [wave][word][lane], FP64 [wave][word][column tile][lane], padding words after each")
        arr("uint64_t", "schur32", self.schur32, hx64, 6)
        arr("uint64_t", "schur64", self.schur64, hx64, 6)
        arr("uint16_t", "diag_pos", self.diag_pos)
        arr("int8_t", "charge", self.charges + [0] * (self.block - self.neq))
        emit_rate_params(L, net)
        L.append(f"}}  // namespace {ns}")
        with open(path, "w") as fh:
            fh.write("\n".join(L) + "\n")

    def _schur_tables(self):
        """Operand positions of the MFMA Schur complement, one u64 word per lane
        for two k steps: A = L[t][k] | B = U[k][j] << 16 of the first step,
        the same for the second << 32.  A k past the head reads LU[luzero] = 0.

        FP32 (v_mfma_f32_32x32x2f32, wave w -> tile rows 32 (w >> 1), columns
        32 (w & 1); lane l = A row / B column l % 32, k offset l / 32): word s
        of wave w holds k = 4s + l/32 and 4s + 2 + l/32; [w][s][lane].
        FP64 (v_mfma_f64_16x16x4f64, wave w -> row band 16w, four column tiles
        jt; lane l = row / column l % 16, k offset l / 16): word s of tile jt
        holds k = 8s + l/16 and 8s + 4 + l/16; [w][s][jt][lane].
        Each table is followed by SCHUR_AHEAD groups of padding words, which
        the loops' look-ahead loads read and never use."""
        kh = self.kh
        z = self.luzero
        lth = lambda t, k: self.lth[t * kh + k] if k < kh else z
        uht = lambda k, j: self.uht[k * 64 + j] if k < kh else z

        def word(t, j, k0, k1):
            return lth(t, k0) | uht(k0, j) << 16 | lth(t, k1) << 32 | uht(k1, j) << 48

        t32, t64 = [], []
        if kh:
            for w in range(4):
                ti, tj = w >> 1, w & 1
                for s in range(kh // 4):
                    for lane in range(64):
                        li, lk = lane & 31, lane >> 5
                        t32.append(word(32 * ti + li, 32 * tj + li, 4 * s + lk, 4 * s + 2 + lk))
            for w in range(4):
                for s in range((kh + 7) // 8):
                    for jt in range(4):
                        for lane in range(64):
                            li, lk = lane & 15, lane >> 4
                            t64.append(word(16 * w + li, 16 * jt + li, 8 * s + lk, 8 * s + 4 + lk))
            pad = z | z << 16 | z << 32 | z << 48
            t32 += [pad] * (SCHUR_AHEAD * SCHUR_G32 * 64)
            t64 += [pad] * (SCHUR_AHEAD * 4 * 64)
        return t32, t64

    def summary(self) -> str:
        ne = sum(self.sp_groups[:self.nep]) * SU
        ni = sum(self.sp_groups[self.nep:]) * SU
        return (f"{self.net.name}: NB={self.nb} FR={self.fr} VL={self.vl} TM={self.tm} PM={self.pm} "
                f"NJ={self.nj} JTM={self.jtm} NLU={self.nlu} (NHV={self.nhv}, NT={self.nt}, KH={self.kh}) "
                f"MTM={self.mtm} levels={self.nlev} elimination phases={self.nep} rounds={ne} "
                f"NHI={self.nhi} inverse phases={len(self.sp_groups) - self.nep} rounds={ni} "
                f"spmv rounds={[sp['rounds'] for sp in self.spmv]}")


def generate(names=("mynetwork", "full"), net_dir=None, out_dir=None) -> List[WgTables]:
    base = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    net_dir = net_dir or os.path.join(base, "networks")
    out_dir = out_dir or os.path.join(base, "csrc", "gen")
    os.makedirs(out_dir, exist_ok=True)
    out = []
    for n in names:
        t = WgTables(load_network(n, net_dir))
        t.emit(os.path.join(out_dir, f"wg_{n}.h"))
        out.append(t)
    return out


if __name__ == "__main__":
    for t in generate():
        print(t.summary())
