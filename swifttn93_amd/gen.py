"""Deterministic synthetic alignments for the benchmark workloads (codes 0..17 = "ACGTURYSWKMBDHVN?-").

Randomness comes from a counter-based splitmix64 stream (numpy uint64 arithmetic), so the bytes are identical on
every machine and numpy version.  Workloads:

  c0   500 x 1,000: uniform random root; 8 clusters on star trees, per-branch substitution rate 0.2-3 %;
       0.5 % IUPAC codes (R, Y, K, M, S, W, N uniform); 2 % gaps in runs of 1-30.
  c1   6,700 x 1,300 HIV-pol-shaped: 200 transmission clusters (sizes geometric, mean 15) at 0.1-1.5 %
       intra-cluster divergence, 5-12 % between clusters (background sequences at the same spread); 1 % ambiguity
       codes in the proportions of the reference's real HIV-1 pol set (application/results/ambiguity_counts.json:
       R 45 %, Y 23 %, N 18 %, M 6 %, W 4 %, K 3 %, S 1 %, and B/D/H/V 0.5 %, the codes whose 1/3 weights make a
       pair order-dependent); 5 % of sequences carry a leading or trailing gap run of 50-400 columns.
  c1s  a smaller sample of the c1 generator (n sequences) for parity tests; `bdhv` raises the share of B/D/H/V
       among the ambiguity codes so the site-order replay is exercised heavily.
  c2   20,000 x 29,903 SARS-CoV-2-shaped: a random recursive genealogy (each sequence copies an earlier one and
       mutates each site with probability 0.001-0.02 %, so depths of about ten edges put sequences 0.01-0.2 % from
       the root); 0-5 N-masked blocks of 100-2,000 columns per sequence; 0.05 % ambiguity codes (c1 proportions).
  c2s  a smaller sample of the c2 generator (n sequences) for parity tests.
  c3   64 alignments of uneven size and length (c0 generator) with mixed -a modes, and c3_bipartite, the 2,000 x
       40,000 query-by-reference block at 1,300 columns (c1 generator).
"""
import numpy as np

ALPHABET = b"ACGTURYSWKMBDHVN?-"
GAP = 17
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class Stream:
    """splitmix64 over a counter: value k of stream `seed` is mix(seed + (k + 1) * golden)."""

    def __init__(self, seed):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.k = 0

    def u64(self, n):
        with np.errstate(over="ignore"):
            c = np.arange(self.k + 1, self.k + 1 + n, dtype=np.uint64)
            z = self.seed + c * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.k += n
        return z

    def uniform(self, n):
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def integers(self, hi, n):
        return (self.uniform(n) * hi).astype(np.int64)


def _mutate(rs, seq, rate):
    """Substitutes each site with probability `rate` by one of the three other bases."""
    L = seq.shape[-1]
    hit = rs.uniform(L) < rate
    shift = rs.integers(3, L) + 1
    out = seq.copy()
    out[hit] = (seq[hit] + shift[hit]) % 4
    return out


def _ambiguities(rs, codes, rate, pool, weights=None):
    n, L = codes.shape
    hit = rs.uniform(n * L).reshape(n, L) < rate
    if weights is None:
        pick = np.array(pool, dtype=np.uint8)[rs.integers(len(pool), n * L).reshape(n, L)]
    else:
        cdf = np.cumsum(np.asarray(weights, dtype=np.float64))
        cdf /= cdf[-1]
        idx = np.minimum(np.searchsorted(cdf, rs.uniform(n * L), side="right"), len(pool) - 1)
        pick = np.array(pool, dtype=np.uint8)[idx].reshape(n, L)
    codes[hit] = pick[hit]


def _gap_runs(rs, codes, frac, lo, hi):
    n, L = codes.shape
    for i in range(n):
        u = rs.uniform(3)
        if u[0] < frac:
            run = int(lo + u[1] * (hi - lo + 1))
            run = min(run, L)
            if u[2] < 0.5:
                codes[i, :run] = GAP
            else:
                codes[i, L - run:] = GAP


def c0(seed=1, N=500, L=1000):
    rs = Stream(seed)
    root = rs.integers(4, L).astype(np.uint8)
    codes = np.empty((N, L), dtype=np.uint8)
    centers = [_mutate(rs, root, 0.03) for _ in range(8)]
    for i in range(N):
        rate = 0.002 + rs.uniform(1)[0] * 0.028
        codes[i] = _mutate(rs, centers[i % 8], rate)
    _ambiguities(rs, codes, 0.005, [5, 6, 9, 10, 7, 8, 15])
    # 2 % gaps in runs of 1-30
    for i in range(N):
        target = int(0.02 * L)
        while target > 0:
            u = rs.uniform(2)
            run = min(1 + int(u[0] * 30), target)
            p = int(u[1] * (L - run))
            codes[i, p:p + run] = GAP
            target -= run
    return codes


# R, Y, N, M, W, K, S, then B, D, H, V
_C1_POOL = [5, 6, 15, 10, 8, 9, 7, 11, 12, 13, 14]
_C1_WEIGHTS = [45.0, 23.0, 18.0, 6.0, 4.0, 3.0, 1.0]


def c1(seed=2, n=6700, L=1300, clusters=200, bdhv=0.005):
    rs = Stream(seed)
    root = rs.integers(4, L).astype(np.uint8)
    codes = np.empty((n, L), dtype=np.uint8)
    i = 0
    # each clustered/background lineage sits at 2.5-6 % from the root, so lineages differ by about 5-12 %
    while i < n:
        center = _mutate(rs, root, 0.025 + 0.035 * rs.uniform(1)[0])
        if clusters > 0:
            clusters -= 1
            size = 1 + int(np.floor(np.log(1.0 - rs.uniform(1)[0]) / np.log(1.0 - 1.0 / 15.0)))
            for _ in range(min(size, n - i)):
                codes[i] = _mutate(rs, center, 0.0005 + 0.007 * rs.uniform(1)[0])
                i += 1
        else:
            codes[i] = center
            i += 1
    # shuffle rows so clusters are spread over the matrix
    perm = np.argsort(rs.u64(n), kind="stable")
    codes = codes[perm]
    w = list(_C1_WEIGHTS)
    w += [sum(w) * bdhv / (1.0 - bdhv) / 4.0] * 4
    _ambiguities(rs, codes, 0.01, _C1_POOL, w)
    _gap_runs(rs, codes, 0.05, 50, 400)
    return codes


def c1s(n=600, seed=3, bdhv=0.07):
    return c1(seed=seed, n=n, L=1300, clusters=max(1, n * 200 // 6700), bdhv=bdhv)


def _poisson(rs, lam):
    """Knuth's product-of-uniforms sampler (lam up to a few hundred)."""
    if lam <= 0:
        return 0
    if lam > 200:  # normal approximation keeps the loop short for large means
        u = rs.uniform(2)
        z = np.sqrt(-2.0 * np.log(max(u[0], 1e-300))) * np.cos(2.0 * np.pi * u[1])
        return max(0, int(round(lam + z * np.sqrt(lam))))
    lim, k, p = np.exp(-lam), 0, 1.0
    while True:
        p *= rs.uniform(1)[0]
        if p <= lim:
            return k
        k += 1


def _mutate_sparse(rs, seq, rate):
    """Like _mutate for small rates: a Poisson number of substituted sites instead of one draw per site."""
    L = seq.shape[-1]
    k = _poisson(rs, L * rate)
    out = seq.copy()
    if k:
        pos = rs.integers(L, k)
        out[pos] = (out[pos] + rs.integers(3, k).astype(np.uint8) + 1) % 4
    return out


def c2(seed=4, n=20000, L=29903):
    rs = Stream(seed)
    codes = np.empty((n, L), dtype=np.uint8)
    codes[0] = rs.integers(4, L).astype(np.uint8)
    for i in range(1, n):
        parent = int(rs.integers(i, 1)[0])
        codes[i] = _mutate_sparse(rs, codes[parent], 0.00001 + 0.00019 * rs.uniform(1)[0])
    perm = np.argsort(rs.u64(n), kind="stable")
    codes = codes[perm]
    # 0-5 N-masked blocks of 100-2,000 columns
    for i in range(n):
        nb = int(rs.integers(6, 1)[0])
        for _ in range(nb):
            u = rs.uniform(2)
            run = 100 + int(u[0] * 1901)
            p = int(u[1] * (L - run))
            codes[i, p:p + run] = 15
    # 0.05 % ambiguity codes in the c1 proportions
    k = _poisson(rs, n * L * 0.0005)
    pos = rs.integers(n * L, k)
    w = list(_C1_WEIGHTS)
    w += [sum(w) * 0.005 / (1.0 - 0.005) / 4.0] * 4
    cdf = np.cumsum(np.asarray(w, dtype=np.float64))
    cdf /= cdf[-1]
    idx = np.minimum(np.searchsorted(cdf, rs.uniform(k), side="right"), len(_C1_POOL) - 1)
    codes.reshape(-1)[pos] = np.array(_C1_POOL, dtype=np.uint8)[idx]
    return codes


def c2s(n=300, seed=5, L=29903):
    return c2(seed=seed, n=n, L=L)


C3_MODES = ["resolve", "average", "skip", "gapmm"]


def c3(seed=6, count=64, n_range=(50, 3000), L_range=(300, 10000)):
    """The batched workload: `count` independent alignments from the c0 generator, sizes log-uniform in 50-3,000
    sequences, lengths uniform in 300-10,000 columns, -a modes cycling resolve/average/skip/gapmm.  A list of
    (codes, mode) pairs.  Parity tests draw from the same generator over smaller ranges."""
    rs = Stream(seed)
    out = []
    n0, n1 = n_range
    l0, l1 = L_range
    for k in range(count):
        u = rs.uniform(3)
        n = int(np.exp(np.log(float(n0)) + u[0] * (np.log(float(n1)) - np.log(float(n0)))))
        L = int(l0 + u[1] * (l1 - l0))
        out.append((c0(seed=int(rs.u64(1)[0] >> np.uint64(1)), N=n, L=L), C3_MODES[k % 4]))
    return out


def c3_bipartite(seed=7, nq=2000, nr=40000, L=1300):
    """The -s bipartite block: nq query and nr reference sequences from the c1 generator (one set, queries first)."""
    n = nq + nr
    return c1(seed=seed, n=n, L=L, clusters=max(1, n * 200 // 6700))


def names_for(n, prefix="s"):
    return [f"{prefix}{i}" for i in range(n)]


def write_fasta(path, codes, names=None):
    codes = np.asarray(codes, dtype=np.uint8)
    names = names or names_for(codes.shape[0])
    lut = np.frombuffer(ALPHABET, dtype=np.uint8)
    with open(path, "wb") as f:
        for nm, row in zip(names, codes):
            f.write(b">" + nm.encode() + b"\n" + lut[row].tobytes() + b"\n")


WORKLOADS = {"c0": c0, "c1": c1, "c1s": c1s, "c2": c2, "c2s": c2s, "c3": c3, "c3_bipartite": c3_bipartite}
