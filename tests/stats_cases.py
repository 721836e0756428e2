"""Oracle-side construction of moment tables and expected p-values."""
import numpy as np


def nbins(size):
    return (int(size) + 99) // 100


def oracle_tables(orc, db, F, S, shard_sizes):
    """Exact compact tables per (file, shard) -> slot arrays for upload."""
    sh, s, e, sid = db
    slot_off = np.full(F * S, -1, np.int64)
    slot_len = np.zeros(F * S, np.uint32)
    means, vars_, tabs = [], [], {}
    off = 0
    for f in range(F):
        for c in range(S):
            if shard_sizes[c] == 0:
                continue
            sel = (sid == f) & (sh == c)
            if not sel.any():
                continue
            nb = nbins(shard_sizes[c])
            _, m, v = orc.chrom_moments_exact(s[sel], e[sel], nb)
            slot_off[f * S + c] = off
            slot_len[f * S + c] = len(m)
            means.append(m)
            vars_.append(v)
            tabs[(f, c)] = (nb, m, v)
            off += len(m)
    mean = np.concatenate(means) if means else np.zeros(0)
    var = np.concatenate(vars_) if vars_ else np.zeros(0)
    return slot_off, slot_len, mean, var, tabs


def oracle_cell_stats(orc, q, r, f, S, shard_sizes, tabs, observed):
    """compute_analytic_stats for query file r vs DB file f (None = no record)."""
    qsh, qs, qe, qf = q
    groups = [[], [], [], [], []]
    tables = []
    slot_of = {}
    for c in range(S):
        if shard_sizes[c] == 0:
            continue
        sel = (qf == r) & (qsh == c)
        if not sel.any():
            continue
        li, ss, ss2, cnt = orc.group_lengths(qe[sel] - qs[sel])
        if (f, c) in tabs:
            slot_of[c] = len(tables)
            tables.append(tabs[(f, c)])
        slot = slot_of.get(c, -1)
        for arr, vals in zip(groups, (li, ss, ss2, cnt, np.full(len(li), slot, np.int32))):
            arr.append(vals)
    if not groups[0]:
        return None
    g = tuple(np.concatenate(a) for a in groups)
    return orc.analytic_stats(g, None, tables, observed)


def mlog10(p):
    return np.inf if p <= 0 else -np.log10(p)
