"""numpy restatement of the replay PPE (test helper, not a test module).

The replay graph's PPE follows the reference's graph-Laplacian rules
(src/lsmps.cpp:29-64): unknowns are the Interior and NearWall rows in index
order, row 0 is the identity, every other row holds -a per unknown neighbor
in (sorted) list order and their sum on the diagonal (1 when not positive),
b = A x* left to right. The edge weight is synthetic: (0.5 + u) / l0^2 with
u the top 53 bits of the splitmix64 finaliser of (lo << 32 | hi) xor
(seed * 0x9E3779B97F4A7C15), lo <= hi the edge's rows; self-edges are
skipped. x* ~ U(-1, 1) is the splitmix64 hash of (xseed, row) used by every
random right-hand side.
"""
import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
# Tool numbering (0 interior, 1 near wall, 2 free surface, 3 splash, 4 wall)
# to ParticleType (0 Interior, 1 FreeSurface, 2 Splash, 3 NearWall, 4 Wall).
TYPE_MAP = np.array([0, 3, 1, 2, 4], np.uint8)


def mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def unit53(h):
    return (np.asarray(h, np.uint64) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def xstar_random(seed, k):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(GOLDEN) + np.asarray(k, np.uint64)
    return -1.0 + 2.0 * unit53(mix64(z))


def replay_weight(seed, lo, hi, l0):
    key = (np.asarray(lo, np.uint64) << np.uint64(32)) | np.asarray(hi, np.uint64)
    with np.errstate(over="ignore"):
        key = key ^ (np.uint64(seed) * np.uint64(GOLDEN))
    return (0.5 + unit53(mix64(key))) / (l0 * l0)


def replay_ppe(row_ptr, col, tool_type, wseed, xseed, l0):
    n = row_ptr.shape[0] - 1
    t = TYPE_MAP[tool_type]
    unknown = (t == 0) | (t == 3)
    rev = np.flatnonzero(unknown)
    pos = -np.ones(n, np.int64)
    pos[rev] = np.arange(rev.size)
    xe = xstar_random(xseed, np.arange(rev.size))
    rows, cols, vals = [0], [], []
    for k, i in enumerate(rev):
        if k == 0:
            cols.append(0)
            vals.append(1.0)
            rows.append(len(cols))
            continue
        nb = np.sort(col[row_ptr[i]:row_ptr[i + 1]])
        d = 0.0
        for j in nb.tolist():
            if j == i or pos[j] < 0:
                continue
            a = float(replay_weight(wseed, min(i, j), max(i, j), l0))
            d += a
            cols.append(int(pos[j]))
            vals.append(-a)
        if d <= 0:
            d = 1.0
        cols.append(k)
        vals.append(d)
        rows.append(len(cols))
    row = np.array(rows, np.int32)
    col_o = np.array(cols, np.int32)
    val = np.array(vals, np.float64)
    b = np.array([sum(val[e] * xe[col_o[e]] for e in range(row[k], row[k + 1])) for k in range(rev.size)])
    return dict(row=row, col=col_o, val=val, b=b, x_exact=xe, rev=rev.astype(np.int32))
