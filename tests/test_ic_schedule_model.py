"""A CPU model of the claim-and-poll protocol of the GPU's IC(0) kernels
(mps_amd/csrc/k_ic.hip: k_ic_levels, k_ic_factor, k_ic_csolve), run before
those kernels go near a GPU and kept as a regression check of the protocol.

The model executes the kernels' state machines literally, one wave
iteration at a time:
  levels  chunks of rows in index order claimed from a ticket, one thread
          per row, batches of 16, levels of the chunk's rows through LDS;
  factor  one wave per row, rows claimed in level order, lane t owning entry
          t: waiting / folding / done stages, batches of 16 of row k's
          entries, the binary search over the row's columns, "only the
          lowest waiting lane polls until its row has come"; and the
          two-phase form (rows of at most 64 entries), whose lanes gather
          their matches into lists in batches of 32 before folding them;
  solves  chunks of the solve order (buckets of levels, index order inside)
          claimed from a ticket, each chunk's rows sorted by level and
          claimed by its waves from an LDS ticket, one wave per row folding
          its products in entry order up to the first value not yet there;
and the wait bounds of all three.  The waves (workgroups for the chunked
kernels) are interleaved by a seeded random scheduler, and every store to
global memory becomes visible to other waves after a random delay (a row's
factor values are flushed before its 1/sqrt(d), as the kernel's s_waitcnt
orders them; nothing else is ordered).  Python floats are IEEE doubles, so
each modelled fold has the kernel's rounding.

Checked here, on PPE matrices of the oracle (the reference's algorithm):
  * every run terminates: no wave waits on a row that no running wave will
    finish, whatever the interleaving and the store delays;
  * the levels equal a serial computation; the factor equals the oracle's
    IC(0) factor bit for bit (3-D sparse and 2-D dense formulas); the solves
    equal a serial restatement of the reference's forward and backward
    substitutions bit for bit;
  * with a wait bound of 1 every wait gives up, the runs still terminate and
    the error flag is raised (the host then reports MPS_ERR_SOLVER).
"""
import random

import numpy as np
import pytest

import harness as H

WAVE = 64
FB = 16
SB = 16


class Memory:
    """Global memory as other waves see it: a store becomes visible after a
    random number of scheduler steps."""

    def __init__(self, rng, max_delay):
        self.rng = rng
        self.max_delay = max_delay
        self.pending = []   # (due step, array, index, value, wave)
        self.now = 0

    def store(self, arr, i, v, wave):
        if self.max_delay == 0:
            arr[i] = v
        else:
            self.pending.append((self.now + self.rng.randint(1, self.max_delay), arr, i, v, wave))

    def flush_wave(self, wave):
        keep = []
        for p in self.pending:
            if p[4] is wave:
                p[1][p[2]] = p[3]
            else:
                keep.append(p)
        self.pending = keep

    def tick(self):
        self.now += 1
        if self.pending:
            keep = []
            for p in self.pending:
                if p[0] <= self.now:
                    p[1][p[2]] = p[3]
                else:
                    keep.append(p)
            self.pending = keep

    def drain(self):
        for p in self.pending:
            p[1][p[2]] = p[3]
        self.pending = []


def run_waves(procs, mem, rng, max_steps=50_000_000):
    live = list(procs)
    steps = 0
    while live:
        w = rng.randrange(len(live))
        try:
            next(live[w])
        except StopIteration:
            live.pop(w)
        mem.tick()
        steps += 1
        assert steps < max_steps, "protocol did not terminate"
    mem.drain()
    return steps


def factor_wave(A, st, mem, dense, wait_limit, ordered=True, order=None):
    """One wave of k_ic_factor (a generator: one yield per wave iteration).
    With ordered=False a row's values may become visible after its
    1/sqrt(d); the readers' re-polling of a not yet visible value covers
    that."""
    rp, col, dpos, aval = A["rp"], A["col"], A["dpos"], A["aval"]
    lval, dinv = st["lval"], st["dinv"]
    me = object()
    while True:
        tk = st["ticket"]
        st["ticket"] += 1
        if tk >= A["n"]:
            return
        e = order[tk] if order is not None else tk
        r0, m = rp[e], dpos[e]
        scol = list(col[r0:r0 + m])
        sval = [None] * m
        pstep = (1 << (m.bit_length() - 1)) if m > 0 else 0
        for g0 in range(0, m, WAVE):
            L = []
            for lane in range(WAVE):
                T = g0 + lane
                act = T < m
                k = scol[T] if act else 0
                a = aval[r0 + T] if act else 0.0
                L.append(dict(T=T, k=k, a=a, kb=rp[k], mk=dpos[k], stage=0 if act else 2, dk=0.0,
                              s=0.0 if dense else a, u=0, bn=0, bi=0, bpos=[-1] * FB, bval=[0.0] * FB, spins=0))
            poll_all = True
            while True:
                if all(x["stage"] == 2 for x in L):
                    break
                donem = [x["stage"] == 2 for x in L]
                waitm = [i for i, x in enumerate(L) if x["stage"] == 0]
                prog = False
                if waitm:
                    lo = waitm[0]
                    for i in waitm:
                        x = L[i]
                        if poll_all or i == lo:
                            d = dinv[x["k"]]
                            if d is not None:
                                x["dk"], x["stage"], x["spins"] = d, 1, 0
                                prog = True
                            else:
                                x["spins"] += 1
                                if x["spins"] > wait_limit:
                                    st["err"] = True
                                    x["stage"] = 1
                    poll_all = L[lo]["stage"] != 0
                for x in L:
                    if x["stage"] == 1 and x["bi"] == x["bn"]:
                        if x["u"] >= x["mk"]:
                            if not dense:
                                v = x["s"] * x["dk"]
                            else:
                                v = 0.0 if abs(x["a"]) < 1.0e-30 else (x["a"] - x["s"]) * (1.0 / x["dk"])
                            sval[x["T"]] = v
                            mem.store(lval, r0 + x["T"], v, me)
                            x["stage"] = 2
                            prog = True
                        else:
                            bn = min(FB, x["mk"] - x["u"])
                            base = x["kb"] + x["u"]
                            bc = [col[base + j] if j < bn else 2**31 - 1 for j in range(FB)]
                            x["bval"] = [lval[base + j] if j < bn else 0.0 for j in range(FB)]
                            ps = [-1] * FB
                            stp = pstep
                            while stp > 0:
                                for j in range(FB):
                                    cand = ps[j] + stp
                                    if cand < x["T"] and scol[cand] < bc[j]:
                                        ps[j] = cand
                                stp >>= 1
                            x["bpos"] = [(ps[j] + 1) if (j < bn and ps[j] + 1 < x["T"] and scol[ps[j] + 1] == bc[j])
                                         else -1 for j in range(FB)]
                            x["bi"], x["bn"] = 0, bn
                            x["u"] += bn
                            prog = True
                for x in L:
                    if x["stage"] == 1 and x["bi"] < x["bn"]:
                        for j in range(FB):
                            if j != x["bi"] or j >= x["bn"]:
                                continue
                            v = x["bpos"][j]
                            if v < 0:
                                x["bi"] += 1
                                prog = True
                                continue
                            lk = x["bval"][j]
                            if lk is None:
                                lk = lval[x["kb"] + x["u"] - x["bn"] + j]
                                x["bval"][j] = lk
                            ready = lk is not None and (v < g0 or donem[v - g0])
                            if ready:
                                lv = sval[v]
                                if not dense:
                                    x["s"] = x["s"] - lv * lk
                                else:
                                    x["s"] = x["s"] + lv * lk
                                x["bi"] += 1
                                prog = True
                                x["spins"] = 0
                            else:
                                x["spins"] += 1
                                if x["spins"] > wait_limit:
                                    st["err"] = True
                                    x["bi"] += 1
                                else:
                                    break
                yield
        akk = aval[r0 + m]
        if not dense:
            d = akk
            for q in range(m):
                d = d - sval[q] * sval[q]
        else:
            sq = 0.0
            for q in range(m):
                sq = sq + sval[q] * sval[q]
            d = akk - sq
        if d <= 0.0:
            d = akk if akk > 0.0 else 1.0e-30
        out = np.sqrt(d) if dense else 1.0 / np.sqrt(d)
        if ordered:
            mem.flush_wave(me)   # s_waitcnt(0): the row's values before its 1/sqrt(d)
        mem.store(dinv, e, float(out), me)
        yield


def solve_wave_group(A, st, mem, back, rin, y, lanes_per_wave, waves, chunk, mode, wait_limit):
    """The waves of one workgroup of k_ic_solve, interleaved inside this
    generator (their LDS is shared); the chunk claim is the workgroup
    barrier.  Positions: A["rows"][t] lists the (column position, value)
    entries of the row at position t, A["dinv"][t] its diagonal factor;
    rin and y are by position.  mode 0 sparse, 1 dense, 2 level."""
    n = A["n"]
    dense = mode == 1
    nch = (n + chunk - 1) // chunk
    sy = [0.0] * chunk
    sflag = [-1] * chunk
    me = object()
    while True:
        tk = st["ticket"]
        st["ticket"] += 1
        if tk >= nch:
            return
        ch = nch - 1 - tk if back else tk
        base, end = ch * chunk, min(ch * chunk + chunk, n)
        rows = []
        for t in range(chunk):
            e = base + t
            r = dict(e=e, t=t, done=e >= n, q=0, bn=0, bi=0, bc=[0] * SB, bl=[0.0] * SB, bv=[0.0] * SB, spins=0)
            if e < n:
                ent = A["rows"][e]
                r["ent"] = ent
                r["m"] = len(ent)
                r["rv"] = rin[e] if mode != 2 else 0.0
                r["s"] = rin[e] if mode == 0 else (0.0 if mode == 1 else -1.0)
            rows.append(r)
        wstate = [dict(poll_all=True, finished=False) for _ in range(waves)]
        while not all(ws["finished"] for ws in wstate):
            order = list(range(waves))
            st["rng"].shuffle(order)
            for w in order:
                ws = wstate[w]
                if ws["finished"]:
                    continue
                lanes = rows[w * lanes_per_wave:(w + 1) * lanes_per_wave]
                if all(r["done"] for r in lanes):
                    ws["finished"] = True
                    continue
                prog = False
                for r in lanes:
                    if not r["done"] and r["bi"] == r["bn"]:
                        if r["q"] >= r["m"]:
                            if mode == 2:
                                out = r["s"] + 1.0
                            else:
                                di = A["dinv"][r["e"]]
                                out = (r["rv"] - r["s"]) / di if dense else r["s"] * di
                            sy[r["t"]] = out
                            sflag[r["t"]] = tk
                            mem.store(y, r["e"], out, me)
                            r["done"] = True
                            prog = True
                        else:
                            bn = min(SB, r["m"] - r["q"])
                            for j in range(bn):
                                r["bc"][j], r["bl"][j] = r["ent"][r["q"] + j]
                            for j in range(SB):
                                c = r["bc"][j]
                                ext = c >= end if back else c < base
                                r["bv"][j] = y[c] if (j < bn and ext) else 0.0
                            r["q"] += bn
                            r["bi"], r["bn"] = 0, bn
                            prog = True
                blocked = {}
                for r in lanes:
                    if r["done"] or r["bi"] >= r["bn"]:
                        continue
                    for j in range(SB):
                        if j != r["bi"] or j >= r["bn"]:
                            continue
                        c = r["bc"][j]
                        ext = c >= end if back else c < base
                        if ext:
                            v = r["bv"][j]
                            ready = v is not None
                        else:
                            ready = sflag[c - base] == tk
                            v = sy[c - base] if ready else 0.0
                        if ready:
                            if mode == 2:
                                r["s"] = max(r["s"], v)
                            elif dense:
                                r["s"] = r["s"] + r["bl"][j] * v
                            else:
                                r["s"] = r["s"] - r["bl"][j] * v
                            r["bi"] += 1
                            prog = True
                            r["spins"] = 0
                        else:
                            r["spins"] += 1
                            if r["spins"] > wait_limit:
                                st["err"] = True
                                r["bi"] += 1
                            else:
                                if ext:
                                    blocked[r["t"]] = j
                                break
                if blocked:
                    lo = min(blocked)
                    got_lo = False
                    for t, j in blocked.items():
                        if ws["poll_all"] or t == lo:
                            r = rows[t]
                            r["bv"][j] = y[r["bc"][j]]
                            if t == lo:
                                got_lo = r["bv"][j] is not None
                    ws["poll_all"] = got_lo
                yield


def csolve_group(A, st, mem, back, rin, y, mode, wait_limit, n_waves, chunk):
    """One workgroup of k_ic_csolve: chunks of `chunk` positions claimed in
    order (BACK: from the end); inside a chunk wave w takes rows w,
    w + n_waves, ... in dependency order.  Per row, lane j holds entry j and
    the value it multiplies -- from the chunk's LDS (visible to the
    workgroup at once) or from global memory (visible after a delay) -- and
    the products are folded in entry order up to the first value not yet
    there, whose lanes poll again.  The waves are interleaved one poll round
    at a time inside this generator."""
    n = A["n"]
    me = object()
    nch = (n + chunk - 1) // chunk
    while True:
        tk = st["ticket"]
        st["ticket"] += 1
        if tk >= nch:
            return
        ch = nch - 1 - tk if back else tk
        cb, ce = ch * chunk, min(ch * chunk + chunk, n)
        sy = {}
        # the chunk's rows by level, then position (BACK: both descending)
        if back:
            srow = sorted(range(cb, ce), key=lambda t: (-A["lev"][t], -t))
        else:
            srow = sorted(range(cb, ce), key=lambda t: (A["lev"][t], t))
        nxt = [0]
        waves = [dict(state=None, done=False) for _ in range(n_waves)]

        def value(c):
            return sy.get(c) if cb <= c < ce else y[c]

        while not all(wv["done"] for wv in waves):
            for wv in st["rng"].sample(waves, len(waves)):
                if wv["done"]:
                    continue
                if wv["state"] is None:
                    if nxt[0] >= len(srow):   # the LDS ticket is exhausted
                        wv["done"] = True
                        continue
                    t = srow[nxt[0]]
                    nxt[0] += 1
                    ent = A["rows"][t]
                    wv["state"] = dict(t=t, ent=ent, g0=0, s=rin[t] if mode == 0 else 0.0, v=None, k=0, spins=0)
                stt = wv["state"]
                ent, g0 = stt["ent"], stt["g0"]
                if g0 < len(ent):
                    cnt = min(WAVE, len(ent) - g0)
                    grp = ent[g0:g0 + cnt]
                    if stt["v"] is None:
                        stt["v"] = [value(c) for c, _ in grp]
                        stt["k"] = 0
                    v = stt["v"]
                    k = stt["k"]
                    while k < cnt and v[k] is not None:
                        pk = grp[k][1] * v[k]
                        stt["s"] = stt["s"] - pk if mode == 0 else stt["s"] + pk
                        k += 1
                    stt["k"] = k
                    if k < cnt:
                        for i in range(k, cnt):
                            if v[i] is None:
                                v[i] = value(grp[i][0])
                        if v[k] is not None:
                            stt["spins"] = 0
                        else:
                            stt["spins"] += 1
                            if stt["spins"] > wait_limit:
                                st["err"] = True
                                v[k] = 0.0
                        continue
                    stt["g0"] += WAVE
                    stt["v"] = None
                    if stt["g0"] < len(ent):
                        continue
                t = stt["t"]
                di = A["dinv"][t]
                out = stt["s"] * di if mode == 0 else (rin[t] - stt["s"]) / di
                sy[t] = out
                mem.store(y, t, out, me)
                wv["state"] = None
            yield


def _matrix(name, **over):
    dim, cfg, pos, vel, typ = H.build(name, solver_type=1, cg_tolerance=1e-8, **over)
    o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1)
    o.step()
    q = o.ppe()
    n = q["n_eq"]
    rp = [int(v) for v in q["row_ptr"]]
    col = [int(v) for v in q["col"]]
    aval = [float(v) for v in q["val"]]
    dpos = []
    for i in range(n):
        p = rp[i]
        while p < rp[i + 1] and col[p] < i:
            p += 1
        dpos.append(p - rp[i])
    return dict(n=n, rp=rp, col=col, aval=aval, dpos=dpos, q=q, dense=dim == 2)


FG2 = 4   # rows of ready lanes loaded together (k_ic_factor2)


def factor2_wave(A, st, mem, dense, wait_limit, ordered=True, order=None):
    """One wave of k_ic_factor2 (rows of at most 64 entries): lane t waits
    for row k; the wave then loads row k whole (the ready lanes' rows FG2 at
    a time, all loaded before any is matched) and puts the matches (position
    v in this row, L_kj) into lane t's list in row k's entry order -- a row
    with a value not yet visible is loaded again next time -- then lane t
    folds its list in order as lanes v finish."""
    rp, col, dpos, aval = A["rp"], A["col"], A["dpos"], A["aval"]
    lval, dinv = st["lval"], st["dinv"]
    me = object()
    while True:
        tk = st["ticket"]
        st["ticket"] += 1
        if tk >= A["n"]:
            return
        e = order[tk] if order is not None else tk
        r0, m = rp[e], dpos[e]
        assert m <= WAVE
        scol = list(col[r0:r0 + m])
        sval = [None] * m
        pstep = (1 << (m.bit_length() - 1)) if m > 0 else 0
        L = []
        for T in range(WAVE):
            act = T < m
            k = scol[T] if act else 0
            L.append(dict(T=T, k=k, a=aval[r0 + T] if act else 0.0, kb=rp[k], mk=dpos[k], stage=0 if act else 3,
                          dk=0.0, s=0.0 if dense else (aval[r0 + T] if act else 0.0), lst=[], mi=0, spins=0))
        poll_all = True
        while True:
            if all(x["stage"] == 3 for x in L):
                break
            waitm = [i for i, x in enumerate(L) if x["stage"] == 0]
            if waitm:
                lo = waitm[0]
                for i in waitm:
                    x = L[i]
                    if poll_all or i == lo:
                        d = dinv[x["k"]]
                        if d is not None:
                            x["dk"], x["stage"], x["spins"] = d, 1, 0
                        else:
                            x["spins"] += 1
                            if x["spins"] > wait_limit:
                                st["err"] = True
                                x["stage"] = 1
                poll_all = L[lo]["stage"] != 0
            ready = [x for x in L if x["stage"] == 1]
            for g0 in range(0, len(ready), FG2):
                grp = ready[g0:g0 + FG2]
                rows = [([col[x["kb"] + j] for j in range(x["mk"])],
                         [lval[x["kb"] + j] for j in range(x["mk"])]) for x in grp]
                for x, (bc, bv) in zip(grp, rows):
                    if any(v is None for v in bv):
                        x["spins"] += 1
                        if x["spins"] <= wait_limit:
                            continue
                        st["err"] = True
                    for j in range(x["mk"]):
                        ps, stp = -1, pstep
                        while stp > 0:
                            cand = ps + stp
                            if cand < x["T"] and scol[cand] < bc[j]:
                                ps = cand
                            stp >>= 1
                        if ps + 1 < x["T"] and scol[ps + 1] == bc[j]:
                            x["lst"].append((ps + 1, 0.0 if bv[j] is None else bv[j]))
                    x["stage"], x["spins"] = 2, 0
            # the fold chain, as many links as are ready
            while True:
                donem = [x["stage"] == 3 for x in L]
                fprog = False
                for x in L:
                    if x["stage"] != 2:
                        continue
                    while x["mi"] < len(x["lst"]):
                        v, lk = x["lst"][x["mi"]]
                        if not donem[v]:
                            break
                        x["s"] = x["s"] - sval[v] * lk if not dense else x["s"] + sval[v] * lk
                        x["mi"] += 1
                        fprog = True
                    if x["mi"] == len(x["lst"]):
                        if not dense:
                            v = x["s"] * x["dk"]
                        else:
                            v = 0.0 if abs(x["a"]) < 1.0e-30 else (x["a"] - x["s"]) * (1.0 / x["dk"])
                        sval[x["T"]] = v
                        mem.store(lval, r0 + x["T"], v, me)
                        x["stage"] = 3
                        fprog = True
                if not fprog:
                    break
            yield
        akk = aval[r0 + m]
        if not dense:
            d = akk
            for q in range(m):
                d = d - sval[q] * sval[q]
        else:
            sq = 0.0
            for q in range(m):
                sq = sq + sval[q] * sval[q]
            d = akk - sq
        if d <= 0.0:
            d = akk if akk > 0.0 else 1.0e-30
        out = np.sqrt(d) if dense else 1.0 / np.sqrt(d)
        if ordered:
            mem.flush_wave(me)
        mem.store(dinv, e, float(out), me)
        yield


def _levels(A):
    """Dependency levels, serially: 1 + the largest level of the row's
    strictly-lower columns, 0 without any."""
    lev = []
    for e in range(A["n"]):
        cols = A["col"][A["rp"][e]:A["rp"][e] + A["dpos"][e]]
        lev.append(1 + max((lev[c] for c in cols), default=-1))
    return lev


def _orders(A, rng):
    """The kernels' level order: rows by level, ties in a random order (any
    order within a level is valid), and its inverse."""
    lev = _levels(A)
    n = A["n"]
    by = {}
    for e in range(n):
        by.setdefault(lev[e], []).append(e)
    order = []
    for L in sorted(by):
        rows = by[L]
        rng.shuffle(rows)
        order += rows
    pos = [0] * n
    for t, e in enumerate(order):
        pos[e] = t
    return lev, order, pos


def _factor(A, seed, n_waves, max_delay, wait_limit, ordered=True, form=1):
    rng = random.Random(seed)
    mem = Memory(rng, max_delay)
    _, fperm, _ = _orders(A, rng)
    st = dict(ticket=0, lval=[None] * len(A["col"]), dinv=[None] * A["n"], err=False)
    wave = factor_wave if form == 1 else factor2_wave
    procs = [wave(A, st, mem, A["dense"], wait_limit, ordered, fperm) for _ in range(n_waves)]
    run_waves(procs, mem, rng)
    return st


def _lower_rows(A, lval, sperm, spos):
    rp, col, dpos = A["rp"], A["col"], A["dpos"]
    return [[(spos[col[rp[e] + q]], lval[rp[e] + q]) for q in range(dpos[e])] for e in sperm]


def _upper_rows(A, lval, dense, sperm=None, spos=None):
    """Row j's L_ij, i > j: ascending i (dense), descending (sparse)."""
    n = A["n"]
    U = [[] for _ in range(n)]
    for i in range(n):   # ascending i
        for q in range(A["dpos"][i]):
            j = A["col"][A["rp"][i] + q]
            U[j].append((i if spos is None else spos[i], lval[A["rp"][i] + q]))
    U = [u if dense else list(reversed(u)) for u in U]
    return U if sperm is None else [U[e] for e in sperm]


def _run_levels(A2, seed, groups, lanes, waves, max_delay, wait_limit):
    rng = random.Random(seed)
    mem = Memory(rng, max_delay)
    st = dict(ticket=0, err=False, rng=rng)
    y = [None] * A2["n"]
    procs = [solve_wave_group(A2, st, mem, False, None, y, lanes, waves, lanes * waves, 2, wait_limit)
             for _ in range(groups)]
    run_waves(procs, mem, rng)
    return y, st["err"]


def _run_csolve(A2, rin, back, mode, seed, groups, max_delay, wait_limit, n_waves=8, chunk=128):
    rng = random.Random(seed)
    mem = Memory(rng, max_delay)
    st = dict(ticket=0, err=False, rng=rng)
    y = [None] * A2["n"]
    procs = [csolve_group(A2, st, mem, back, rin, y, mode, wait_limit, n_waves, chunk) for _ in range(groups)]
    run_waves(procs, mem, rng)
    return y, st["err"]


def _solves(A, lval, dinv, r, seed, groups, max_delay, wait_limit, n_waves=8, chunk=128):
    """Forward then backward substitution by position of the solve order
    (buckets of levels, index order inside); results returned in eq order."""
    rng = random.Random(seed)
    dense = A["dense"]
    lev, order, pos = _solve_order(A)
    n = A["n"]
    dp = [dinv[e] for e in order]
    rin = [r[e] for e in order]
    mode = 1 if dense else 0
    levp = [lev[e] for e in order]
    AL = dict(n=n, rows=_lower_rows(A, lval, order, pos), dinv=dp, lev=levp)
    yp, e1 = _run_csolve(AL, rin, False, mode, seed + 1, groups, max_delay, wait_limit, n_waves, chunk)
    AU = dict(n=n, rows=_upper_rows(A, lval, dense, order, pos), dinv=dp, lev=levp)
    zp, e2 = _run_csolve(AU, yp, True, mode, seed + 2, groups, max_delay, wait_limit, n_waves, chunk)
    y = [None] * n
    z = [None] * n
    for t, e in enumerate(order):
        y[e], z[e] = yp[t], zp[t]
    return y, z, e1 or e2


def _solve_order(A, bucket_min=32):
    """The solve order of ic_order: rows by bucket of 2^g levels (g the
    smallest with at least bucket_min levels and at most 256 buckets),
    ascending index inside a bucket, and its inverse."""
    lev = _levels(A)
    n = A["n"]
    nlev = max(lev) + 1 if n else 1
    g = 0
    while (1 << g) < bucket_min:
        g += 1
    while ((nlev - 1) >> g) >= 256:
        g += 1
    order = sorted(range(n), key=lambda e: (lev[e] >> g, e))
    pos = [0] * n
    for t, e in enumerate(order):
        pos[e] = t
    return lev, order, pos


def _serial_solves(A, lval, dinv, r, dense):
    """forward_solve_sparse / backward_solve_sparse (3DP pressure_solver.c:
    235-255; the dense forms of mps_fluid2D/src/pressure_solver.c:120-149),
    run serially in the reference's order."""
    n = A["n"]
    rp, col, dpos = A["rp"], A["col"], A["dpos"]
    y = [0.0] * n
    for i in range(n):
        if not dense:
            s = r[i]
            for q in range(dpos[i]):
                s = s - lval[rp[i] + q] * y[col[rp[i] + q]]
            y[i] = s * dinv[i]
        else:
            s = 0.0
            for q in range(dpos[i]):
                s = s + lval[rp[i] + q] * y[col[rp[i] + q]]
            y[i] = (r[i] - s) / dinv[i]
    if not dense:
        z = list(y)
        for i in range(n - 1, -1, -1):
            z[i] = z[i] * dinv[i]
            for q in range(dpos[i]):
                j = col[rp[i] + q]
                z[j] = z[j] - lval[rp[i] + q] * z[i]
        return y, z
    z = [0.0] * n
    U = _upper_rows(A, lval, True)
    for j in range(n - 1, -1, -1):
        s = 0.0
        for i, l in U[j]:
            s = s + l * z[i]
        z[j] = (y[j] - s) / dinv[j]
    return y, z


FACTOR_CASES = [("dam2d_l0.01", 1, 8, 0, True), ("dam2d_l0.01", 2, 24, 40, True),
                ("dam2d_l0.01", 3, 3, 200, True), ("dam2d_l0.01", 9, 12, 100, False),
                ("fric3d_tiny", 2, 24, 40, True), ("fric3d_tiny", 9, 12, 100, False)]
LEVEL_CASES = [("dam2d_l0.01", 4, 6, 4, 4, 0), ("dam2d_l0.01", 5, 20, 8, 2, 30),
               ("dam2d_l0.01", 6, 3, 16, 4, 300), ("fric3d_tiny", 6, 3, 16, 4, 300)]
SOLVE_CASES = [("dam2d_l0.01", 4, 4, 0), ("dam2d_l0.01", 5, 12, 30), ("dam2d_l0.01", 6, 2, 300),
               ("fric3d_tiny", 7, 6, 100)]


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("name,seed,n_waves,max_delay,ordered", FACTOR_CASES)
def test_factor_protocol_terminates_and_matches_oracle(name, seed, n_waves, max_delay, ordered, form):
    """Both forms of the factor: one-phase (k_ic_factor) and two-phase
    (k_ic_factor2, match lists first)."""
    A = _matrix(name)
    st = _factor(A, seed, n_waves, max_delay, wait_limit=10**12, ordered=ordered, form=form)
    assert not st["err"]
    q = A["q"]
    lv = [st["lval"][A["rp"][e] + t] for e in range(A["n"]) for t in range(A["dpos"][e])]
    H.say(name, "n_eq", A["n"], "lower entries", len(lv))
    assert np.array_equal(np.array(lv).view(np.uint64), q["l_val"].view(np.uint64))
    assert np.array_equal(np.array(st["dinv"]).view(np.uint64), q["l_diag"].view(np.uint64))


@pytest.mark.parametrize("name,seed,groups,max_delay", SOLVE_CASES)
def test_solve_protocol_terminates_and_matches_serial(name, seed, groups, max_delay):
    A = _matrix(name)
    st = _factor(A, 11, 8, 0, wait_limit=10**12)
    rng = np.random.default_rng(seed)
    r = [float(v) for v in rng.standard_normal(A["n"])]
    y, z, err = _solves(A, st["lval"], st["dinv"], r, seed, groups, max_delay, 10**12)
    assert not err
    ys, zs = _serial_solves(A, st["lval"], st["dinv"], r, A["dense"])
    assert np.array_equal(np.array(y).view(np.uint64), np.array(ys).view(np.uint64))
    assert np.array_equal(np.array(z).view(np.uint64), np.array(zs).view(np.uint64))


@pytest.mark.parametrize("name,seed,groups,lanes,waves,max_delay", LEVEL_CASES)
def test_level_pass_protocol(name, seed, groups, lanes, waves, max_delay):
    """The level pass (k_ic_solve mode 2 over the eq-order CSR, chunks in
    index order) gives the serial levels, under any interleaving."""
    A = _matrix(name)
    rows = [[(A["col"][A["rp"][e] + q], None) for q in range(A["dpos"][e])] for e in range(A["n"])]
    lev, err = _run_levels(dict(n=A["n"], rows=rows, dinv=None), seed, groups, lanes, waves, max_delay, 10**12)
    assert not err
    assert [int(v) for v in lev] == _levels(A)
    H.say(name, "levels", max(lev) + 1)


def test_wait_bound_gives_up_and_terminates():
    A = _matrix("fric3d_tiny")
    for form in (1, 2):
        st = _factor(A, 7, 16, 50, wait_limit=1, form=form)
        assert st["err"]
        assert all(d is not None for d in st["dinv"])
    rng = np.random.default_rng(0)
    r = [float(v) for v in rng.standard_normal(A["n"])]
    st2 = _factor(A, 11, 8, 0, wait_limit=10**12)
    y, z, err = _solves(A, st2["lval"], st2["dinv"], r, 8, 6, 50, 1)
    assert err
    assert all(v is not None for v in y) and all(v is not None for v in z)
