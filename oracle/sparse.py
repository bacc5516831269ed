"""CPU restatement of voltax/sparse.py (oracle / test infrastructure).

* `structure`            CSR pattern + stamp positions       voltax/sparse.py:73-193
* `jacobian`             stamped ``f*df/dz + q*dq/dz``        voltax/sparse.py:200-249
* `pivot_groups`         union-find + bipartite matching      voltax/sparse.py:257-305
* `nested_dissection`    BFS-level separators, postorder      voltax/sparse.py:308-378
* `make_plan`            supernodes, borders, level buckets   voltax/sparse.py:381-535
* `factor` / `solve`     block LDU with inverted pivots       voltax/sparse.py:543-595
* `spsolve_vjp`          adjoint of ``x = A^-1 b``            voltax/sparse.py:598-621

Everything is FP64 (or complex128) numpy; the scipy graph routines are the
ones the reference calls (scipy 1.15.3 in the reference lock file).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

from . import dual as D
from .circuit import Circuit, charges, currents

LEAF_SIZE = 16
SPLIT_SAVING = 1e6


@dataclass
class Structure:
    size: int
    n_nodes: int
    indptr: np.ndarray
    indices: np.ndarray
    stamps: dict  # group -> (N, m, m) positions (nnz = dropped)
    devices: dict  # group -> (N, m) unknowns, -1 = ground
    terminals: dict  # group -> T

    @property
    def nnz(self):
        return len(self.indices)

    @property
    def rows(self):
        return np.repeat(np.arange(self.size), np.diff(self.indptr))

    def positions(self, r, c):
        """Slot of ``(r, c)`` in the value array; ``nnz`` if either is ground."""
        r, c = np.broadcast_arrays(np.asarray(r), np.asarray(c))
        keys = self.rows.astype(np.int64) * self.size + self.indices
        want = r.astype(np.int64) * self.size + c
        pos = np.searchsorted(keys, want)
        ok = (r >= 0) & (c >= 0)
        hit = np.where(ok, keys[np.minimum(pos, len(keys) - 1)] == want, True)
        assert hit.all(), "entry outside the sparsity pattern"
        return np.where(ok, pos, len(keys))

    @property
    def diagonal(self):
        a = np.arange(self.size)
        return self.positions(a, a)

    def to_dense(self, data):
        A = np.zeros((self.size, self.size), np.asarray(data).dtype)
        A[self.rows, self.indices] = data
        return A

    def matvec(self, data, x):
        return np.bincount(self.rows, weights=data * x[self.indices],
                           minlength=self.size) if not np.iscomplexobj(data) else (
            self.to_dense(data) @ x)


def structure(c: Circuit) -> Structure:
    S = c.size
    rows, cols, devs, terms = [np.arange(S)], [np.arange(S)], {}, {}
    for name in c.groups:
        u = c.unknowns(name)
        if u.shape[0] == 0:
            continue
        devs[name] = u
        terms[name] = c.groups[name].n_terminals
        m = u.shape[1]
        r = np.repeat(u[:, :, None], m, 2).ravel()
        cc = np.repeat(u[:, None, :], m, 1).ravel()
        keep = (r >= 0) & (cc >= 0)
        rows.append(r[keep])
        cols.append(cc[keep])
    r, cc = np.concatenate(rows), np.concatenate(cols)
    pat = sp.csr_matrix((np.ones(len(r)), (r, cc)), shape=(S, S))
    pat.sum_duplicates()
    pat.sort_indices()
    st = Structure(S, c.n_nodes, pat.indptr.astype(np.int64),
                   pat.indices.astype(np.int64), {}, devs, terms)
    for name, u in devs.items():
        st.stamps[name] = st.positions(u[:, :, None], u[:, None, :])
    return st


def local_blocks(g, u0, t, cf, cq, params=None):
    """Per-device ``cf * d(I,F)/du + cq * d(Q,P)/du`` blocks ``(N, m, m)`` and
    the local values ``cf*(I,F) + cq*(Q,P)`` ``(N, m)``."""
    n, m = u0.shape
    T = g.n_terminals
    p = g.params if params is None else params
    useq = g.has_charges and cq != 0
    us = [D.Dual.seed(u0[:, j], j, m) for j in range(m)]
    vals = np.zeros((n, m))
    blk = np.zeros((n, m, m))

    def acc(A, B, coef):
        for off, lst in ((0, A), (T, B)):
            for j, a in enumerate(lst or []):
                vals[:, off + j] += coef * np.broadcast_to(D.val(a), (n,))
                blk[:, off + j, :] += coef * D.tan(a, (n, m))

    if cf != 0:
        acc(*currents(g, us[:T], us[T:], t, p), cf)
    if useq:
        acc(*charges(g, us[:T], us[T:], p), cq)
    return blk, vals


def jacobian(c: Circuit, z, t=0.0, f=1.0, q=0.0, st: Structure | None = None):
    st = structure(c) if st is None else st
    zext = np.append(z, 0.0)
    data = np.zeros(st.nnz + 1)
    for name, u in st.devices.items():
        g = c.groups[name]
        cq = q if g.has_charges else 0
        if f == 0 and cq == 0:
            continue
        blk, _ = local_blocks(g, zext[u], t, f, cq)
        np.add.at(data, st.stamps[name].ravel(), blk.ravel())
    return data[: st.nnz]


# =============================================================================
# Symbolic analysis
# =============================================================================


def pivot_groups(st: Structure) -> np.ndarray:
    """Group id per unknown: zero-diagonal unknowns glued to partners."""
    S = st.size
    parent = np.arange(S)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[rb] = ra

    internals, cands, whole = [], [], []
    for name, u in st.devices.items():
        T = st.terminals[name]
        k = u.shape[1] - T
        if k == 0:
            continue
        for dev in u:
            terms = [int(v) for v in dev[:T] if v >= 0]
            ints = [int(v) for v in dev[T:]]
            if T == 2 and k == 1 and terms:
                internals.append(ints[0])
                cands.append(terms)
            else:
                whole.append(terms + ints)
    if internals:
        r = np.repeat(np.arange(len(internals)), [len(x) for x in cands])
        cc = np.concatenate(cands)
        bip = sp.csr_matrix((np.ones(len(r)), (r, cc)),
                            shape=(len(internals), st.n_nodes))
        match = csg.maximum_bipartite_matching(bip, perm_type="column")
        for i, (x, cl) in enumerate(zip(internals, cands)):
            if match[i] >= 0:
                union(int(match[i]), x)
            else:
                whole.append(cl + [x])
    for grp in whole:
        for v in grp[1:]:
            union(grp[0], v)
    roots = np.array([find(a) for a in range(S)])
    return np.unique(roots, return_inverse=True)[1]


def nested_dissection(adj, w, leaf, last):
    """Supernode vertex sets in postorder and their parent indices."""
    found = []
    top = -1
    if last.any():
        found.append((np.flatnonzero(last), -1))
        top = 0
    todo = [(np.flatnonzero(~last), top)]
    while todo:
        vs, par = todo.pop()
        if len(vs) == 0:
            continue
        sub = adj[vs][:, vs]
        nc, lab = csg.connected_components(sub, directed=False)
        if nc > 1:
            todo += [(vs[lab == i], par) for i in range(nc)]
            continue
        if w[vs].sum() <= leaf:
            found.append((vs, par))
            continue
        start = 0
        for _ in range(4):  # pseudo-peripheral start vertex
            dist = csg.shortest_path(sub, unweighted=True, indices=start)
            far = int(np.argmax(dist))
            if dist[far] <= dist[start] or far == start:
                break
            start = far
        dist = csg.shortest_path(sub, unweighted=True, indices=start).astype(int)
        depth = dist.max()
        if depth < 2:
            found.append((vs, par))
            continue
        lw = np.bincount(dist, weights=w[vs], minlength=depth + 1)
        cum = np.cumsum(lw)
        tot = cum[-1]
        ok = [lv for lv in range(1, depth)
              if cum[lv - 1] >= tot / 3 and tot - cum[lv] >= tot / 3]
        if not ok:
            mid = int(np.searchsorted(cum, tot / 2))
            ok = [min(max(mid, 1), depth - 1)]
        lv = min(ok, key=lambda c_: lw[c_])
        touch = (sub @ (dist == lv + 1).astype(float)) > 0
        sep = (dist == lv) & touch
        me = len(found)
        found.append((vs[sep], par))
        todo.append((vs[(dist < lv) | ((dist == lv) & ~touch)], me))
        todo.append((vs[dist > lv], me))
    kids = [[] for _ in found]
    roots = []
    for i, (_, p) in enumerate(found):
        (kids[p] if p >= 0 else roots).append(i)
    order, stack = [], [(r, False) for r in reversed(roots)]
    while stack:
        i, done = stack.pop()
        if done:
            order.append(i)
        else:
            stack.append((i, True))
            stack += [(ch, False) for ch in reversed(kids[i])]
    new = {old: i for i, old in enumerate(order)}
    sets = [found[o][0] for o in order]
    parents = np.array([new[found[o][1]] if found[o][1] >= 0 else -1 for o in order],
                       np.int64)
    return sets, parents


@dataclass
class Bucket:
    m: int
    kp: int
    rp: int
    off: int
    piv: np.ndarray  # (m, kp), S = padding
    border: np.ndarray  # (m, rp), S = padding
    valid: np.ndarray  # (m, kp)
    schur: np.ndarray  # (m, rp, rp) positions in work, work = dropped


@dataclass
class Plan:
    size: int
    work: int
    buckets: list
    a_pos: np.ndarray
    stats: dict
    supernodes: list
    parents: np.ndarray
    borders: list


def make_plan(st: Structure, leaf: int = LEAF_SIZE) -> Plan:
    S = st.size
    gid = pivot_groups(st)
    G = int(gid.max()) + 1
    E = sp.csr_matrix((np.ones(S), (np.arange(S), gid)), shape=(S, G))
    P = sp.csr_matrix((np.ones(st.nnz), st.indices, st.indptr), shape=(S, S))
    Ps = (P + P.T).tocsr()
    Q = (E.T @ Ps @ E).tocsr()
    Q.setdiag(0)
    Q.eliminate_zeros()
    w = np.bincount(gid, minlength=G).astype(float)
    deg = np.diff(Q.indptr)
    last = np.ones(G, bool) if S <= 2 * leaf else deg > max(32.0, 8.0 * np.sqrt(G))
    qsets, parents = nested_dissection(Q, w, leaf, last)
    order = np.argsort(gid, kind="stable")
    starts = np.searchsorted(gid[order], np.arange(G + 1))
    sns = [np.sort(np.concatenate([order[starts[g]:starts[g + 1]] for g in qs]))
           for qs in qsets]
    nsn = len(sns)
    sn_of = np.empty(S, np.int64)
    pos_in = np.empty(S, np.int64)
    for i, s in enumerate(sns):
        sn_of[s] = i
        pos_in[s] = np.arange(len(s))
    kids = [[] for _ in range(nsn)]
    for i, p in enumerate(parents):
        if p >= 0:
            kids[p].append(i)
    borders, level = [], np.zeros(nsn, np.int64)
    for i, s in enumerate(sns):
        nb = Ps.indices[np.concatenate([np.arange(Ps.indptr[v], Ps.indptr[v + 1])
                                        for v in s])]
        cand = np.unique(np.concatenate([nb] + [borders[c] for c in kids[i]]))
        borders.append(cand[sn_of[cand] > i])
        if kids[i]:
            level[i] = 1 + max(level[c] for c in kids[i])
    k = np.array([len(s) for s in sns])
    r = np.array([len(b) for b in borders])

    def cost(sel):
        if len(sel) == 0:
            return 0.0
        K, R = k[sel].max(), r[sel].max()
        return len(sel) * (2.0 * K ** 3 + 2.0 * K * K * R + 2.0 * K * R * R + 64.0)

    classes = []
    for lv in range(int(level.max()) + 1):
        sel = np.flatnonzero(level == lv)
        sel = sel[np.argsort(k[sel] ** 3 + k[sel] ** 2 * r[sel] + k[sel] * r[sel] ** 2,
                             kind="stable")]
        cut = min(range(len(sel)), key=lambda c_: cost(sel[:c_]) + cost(sel[c_:]))
        if cost(sel) - cost(sel[:cut]) - cost(sel[cut:]) < SPLIT_SAVING:
            cut = 0
        classes += [part for part in (sel[:cut], sel[cut:]) if len(part)]
    off = 0
    b11, b12, b21 = (np.zeros(nsn, np.int64) for _ in range(3))
    KP, RP = np.zeros(nsn, np.int64), np.zeros(nsn, np.int64)
    lay = []
    for sel in classes:
        m, kp, rp = len(sel), int(k[sel].max()), int(r[sel].max())
        j = np.arange(m)
        b11[sel] = off + j * kp * kp
        b12[sel] = off + m * kp * kp + j * kp * rp
        b21[sel] = off + m * kp * kp + m * kp * rp + j * rp * kp
        KP[sel], RP[sel] = kp, rp
        lay.append((sel, m, kp, rp, off))
        off += m * (kp * kp + 2 * kp * rp)
    work = off
    bsn = np.repeat(np.arange(nsn), r)
    bvar = np.concatenate(borders) if nsn else np.zeros(0, np.int64)
    bkeys = bsn * S + bvar
    bo = np.argsort(bkeys)
    bkeys = bkeys[bo]
    bidx = np.concatenate([np.arange(x) for x in r])[bo] if nsn else bkeys

    def bpos(o, v):
        if len(bidx) == 0:
            return np.zeros(np.shape(o), np.int64)
        qq = np.searchsorted(bkeys, o * S + v)
        return bidx[np.minimum(qq, len(bidx) - 1)]

    def position(u, v):
        su, sv = sn_of[u], sn_of[v]
        o = np.minimum(su, sv)
        same = b11[o] + pos_in[u] * KP[o] + pos_in[v]
        upper = b12[o] + pos_in[u] * RP[o] + bpos(o, v)
        lower = b21[o] + bpos(o, u) * KP[o] + pos_in[v]
        return np.where(su == sv, same, np.where(su < sv, upper, lower))

    a_pos = position(st.rows, st.indices)
    buckets = []
    for sel, m, kp, rp, boff in lay:
        piv = np.full((m, kp), S, np.int64)
        brd = np.full((m, rp), S, np.int64)
        schur = np.full((m, rp, rp), work, np.int64)
        for j, i in enumerate(sel):
            piv[j, : k[i]] = sns[i]
            brd[j, : r[i]] = borders[i]
            if r[i]:
                b = borders[i]
                schur[j, : r[i], : r[i]] = position(b[:, None], b[None, :])
        valid = np.arange(kp)[None, :] < k[sel][:, None]
        buckets.append(Bucket(m, kp, rp, boff, piv, brd, valid, schur))
    stats = dict(supernodes=nsn, levels=int(level.max()) + 1, buckets=len(buckets),
                 factor_entries=int(np.sum(k * k + 2 * k * r)), work=work,
                 flops=float(np.sum(2.0 * k ** 3 + 2.0 * k * k * r + 2.0 * k * r * r)),
                 lu_flops=float(np.sum(2.0 / 3 * k ** 3 + 2.0 * k * k * r
                                       + 2.0 * k * r * r)),
                 largest_supernode=int(k.max()), pivot_groups=G)
    return Plan(S, work, buckets, a_pos, stats, sns, parents, borders)


# =============================================================================
# Numeric phase
# =============================================================================


def factor(plan: Plan, data):
    work = np.zeros(plan.work + 1, np.asarray(data).dtype)
    np.add.at(work, plan.a_pos, data)
    out = []
    for bk in plan.buckets:
        m, kp, rp, o = bk.m, bk.kp, bk.rp, bk.off
        F11 = work[o: o + m * kp * kp].reshape(m, kp, kp).copy()
        F11 += (~bk.valid)[:, :, None] * np.eye(kp)
        W = np.linalg.inv(F11)
        if rp:
            o2 = o + m * kp * kp
            F12 = work[o2: o2 + m * kp * rp].reshape(m, kp, rp)
            o3 = o2 + m * kp * rp
            F21 = work[o3: o3 + m * rp * kp].reshape(m, rp, kp)
            Ub, Lb = W @ F12, F21 @ W
            np.add.at(work, bk.schur.ravel(), -(F21 @ Ub).ravel())
            work[plan.work] = 0.0
        else:
            Ub = Lb = None
        out.append((W, Ub, Lb))
    return out


def solve(plan: Plan, factors, b, transpose=False):
    S = plan.size
    x = np.append(np.array(b, dtype=np.result_type(b, factors[0][0])), 0.0)
    pairs = list(zip(plan.buckets, factors))
    for bk, (W, Ub, Lb) in pairs:
        y = x[bk.piv]
        y[bk.piv == S] = 0
        if bk.rp:
            off = Ub if transpose else np.swapaxes(Lb, 1, 2)
            np.add.at(x, bk.border.ravel(), -np.einsum("mkr,mk->mr", off, y).ravel())
        Wop = np.swapaxes(W, 1, 2) if transpose else W
        x[bk.piv.ravel()] = np.einsum("mjk,mk->mj", Wop, y).ravel()
        x[S] = 0
    for bk, (W, Ub, Lb) in reversed(pairs):
        if bk.rp:
            off = np.swapaxes(Lb, 1, 2) if transpose else Ub
            xb = x[bk.border]
            xb[bk.border == S] = 0
            np.add.at(x, bk.piv.ravel(), -np.einsum("mkr,mr->mk", off, xb).ravel())
            x[S] = 0
    return x[:S]


def spsolve_vjp(st: Structure, plan: Plan, data, b, g):
    """Cotangents ``(g_data, g_b)`` of ``x = A^-1 b`` (voltax/sparse.py:614-618)."""
    fac = factor(plan, data)
    x = solve(plan, fac, b)
    gb = solve(plan, fac, g, transpose=True)
    return -gb[st.rows] * x[st.indices], gb
