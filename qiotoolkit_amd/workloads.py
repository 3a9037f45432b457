"""Seeded synthetic problems in the reference's JSON term format
(``{"c": coefficient, "ids": [variable ids]}``, cpp/graph/cost_edge.h:26,
cpp/graph/edge.cc:43).

* ``lattice2d``   square L x L periodic Ising, J in {-1,+1}, h = 0
* ``lattice3d``   cubic L^3 periodic Ising spin glass, J ~ N(0,1) rounded to fp32
* ``random_pubo`` PUBO with given term-degree mix, ids without replacement per
                  term, integer coefficients uniform in [-cmax, cmax] \\ {0}
* ``er_ising``    Erdos-Renyi Ising, J uniform int in [-jmax, jmax] \\ {0}, every
                  variable gets a one-variable field term (h uniform in {-1,0,1},
                  zero kept so no variable vanishes)
* ``er_ising_f32`` Erdos-Renyi Ising with J ~ N(0,1) fp32, h = 0
"""
import numpy as np


def _f32(x):
    return float(np.float32(x))


def lattice2d(L=32, seed=0):
    rng = np.random.default_rng(seed)
    terms = []
    for y in range(L):
        for x in range(L):
            i = x + L * y
            right = (x + 1) % L + L * y
            down = x + L * ((y + 1) % L)
            terms.append({"c": int(rng.choice([-1, 1])), "ids": [i, right]})
            terms.append({"c": int(rng.choice([-1, 1])), "ids": [i, down]})
    return terms


def lattice3d(L=16, seed=0):
    rng = np.random.default_rng(seed)
    terms = []
    for z in range(L):
        for y in range(L):
            for x in range(L):
                i = x + L * y + L * L * z
                for nb in ((x + 1) % L + L * y + L * L * z,
                           x + L * ((y + 1) % L) + L * L * z,
                           x + L * y + L * L * ((z + 1) % L)):
                    terms.append({"c": _f32(rng.standard_normal()), "ids": [i, nb]})
    return terms


def random_pubo(n_vars=2000, n_terms=20000, fractions=(0.1, 0.4, 0.3, 0.2), cmax=10, seed=0):
    rng = np.random.default_rng(seed)
    counts = [int(round(f * n_terms)) for f in fractions]
    counts[-1] = n_terms - sum(counts[:-1])
    vals = np.array([v for v in range(-cmax, cmax + 1) if v != 0])
    terms = []
    for deg, cnt in enumerate(counts, start=1):
        for _ in range(cnt):
            ids = rng.choice(n_vars, size=deg, replace=False)
            terms.append({"c": int(rng.choice(vals)), "ids": [int(v) for v in ids]})
    order = rng.permutation(len(terms))
    return [terms[k] for k in order]


def er_ising(N=256, mean_degree=6.0, jmax=3, seed=0, field=True):
    rng = np.random.default_rng(seed)
    p = min(1.0, mean_degree / max(1, N - 1))
    terms = []
    if field:
        for i in range(N):
            terms.append({"c": int(rng.integers(-1, 2)), "ids": [i]})
    vals = np.array([v for v in range(-jmax, jmax + 1) if v != 0])
    iu, ju = np.triu_indices(N, k=1)
    mask = rng.random(len(iu)) < p
    for i, j in zip(iu[mask], ju[mask]):
        terms.append({"c": int(rng.choice(vals)), "ids": [int(i), int(j)]})
    return terms


def er_ising_f32(N=256, mean_degree=8.0, seed=0):
    rng = np.random.default_rng(seed)
    p = min(1.0, mean_degree / max(1, N - 1))
    # sample edges without the O(N^2) mask for large N
    m = rng.binomial(N * (N - 1) // 2, p)
    seen = set()
    terms = []
    while len(terms) < m:
        i, j = rng.integers(0, N, size=2)
        if i == j:
            continue
        a, b = (int(i), int(j)) if i < j else (int(j), int(i))
        if (a, b) in seen:
            continue
        seen.add((a, b))
        terms.append({"c": _f32(rng.standard_normal()), "ids": [a, b]})
    # make sure every variable appears
    present = np.zeros(N, dtype=bool)
    for t in terms:
        present[t["ids"]] = True
    for i in np.nonzero(~present)[0]:
        terms.append({"c": 0.0, "ids": [int(i)]})
    return terms


def log_uniform_sizes(n, lo, hi, seed=0):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)


def geometric_temperatures(t0=0.2, t1=2.0, count=32):
    """Geometric schedule discretised to `count` values (schedule/geometric.cc:9-12)."""
    return [t0 * (t1 / t0) ** (i / (count - 1)) for i in range(count)]


def er_ising_arrays(N, mean_degree=6.0, jmax=3, seed=0):
    """Erdos-Renyi Ising as a term CSR (offsets, ids, coeffs) in O(edges):
    N one-variable field terms h in {-1,0,1} (zero kept), then G(N, p) edges
    with J uniform in [-jmax, jmax] \\ {0}."""
    rng = np.random.default_rng(seed)
    p = min(1.0, mean_degree / max(1, N - 1))
    pairs = N * (N - 1) // 2
    m = int(rng.binomial(pairs, p))
    if m * 4 > pairs:
        iu, ju = np.triu_indices(N, k=1)
        pick = rng.choice(len(iu), size=m, replace=False)
        a, b = iu[pick], ju[pick]
    else:
        # distinct unordered pairs, encoded a*N + b (a < b)
        keys = np.zeros(0, dtype=np.int64)
        while len(keys) < m:
            x = rng.integers(0, N, size=(2 * (m - len(keys)) + 16, 2))
            x = x[x[:, 0] != x[:, 1]]
            lo, hi = np.minimum(x[:, 0], x[:, 1]), np.maximum(x[:, 0], x[:, 1])
            keys = np.unique(np.concatenate([keys, lo * N + hi]))
        keys = keys[rng.permutation(len(keys))[:m]]
        a, b = keys // N, keys % N
    vals = np.array([v for v in range(-jmax, jmax + 1) if v != 0])
    h = rng.integers(-1, 2, size=N)
    J = rng.choice(vals, size=m)
    offsets = np.concatenate([np.arange(N + 1), N + 2 * np.arange(1, m + 1)]).astype(np.int64)
    ids = np.concatenate([np.arange(N), np.stack([a, b], axis=1).ravel()]).astype(np.int32)
    coeffs = np.concatenate([h, J]).astype(np.float64)
    return offsets, ids, coeffs


def _distinct_pairs(rng, N, m):
    """m distinct unordered pairs (a < b) of 0..N-1, uniformly."""
    pairs = N * (N - 1) // 2
    if m * 4 > pairs:
        iu, ju = np.triu_indices(N, k=1)
        pick = rng.choice(len(iu), size=m, replace=False)
        return iu[pick], ju[pick]
    keys = np.zeros(0, dtype=np.int64)
    while len(keys) < m:
        x = rng.integers(0, N, size=(2 * (m - len(keys)) + 16, 2))
        x = x[x[:, 0] != x[:, 1]]
        lo, hi = np.minimum(x[:, 0], x[:, 1]), np.maximum(x[:, 0], x[:, 1])
        keys = np.unique(np.concatenate([keys, lo * N + hi]))
    keys = keys[rng.permutation(len(keys))[:m]]
    return keys // N, keys % N


def er_ising_f32_arrays(N, mean_degree=8.0, seed=0):
    """Erdos-Renyi Ising with J ~ N(0,1) rounded to fp32, h = 0, as a term
    CSR; variables without an edge get a zero one-variable term so that all
    N exist."""
    rng = np.random.default_rng(seed)
    p = min(1.0, mean_degree / max(1, N - 1))
    m = int(rng.binomial(N * (N - 1) // 2, p))
    a, b = _distinct_pairs(rng, N, m)
    J = rng.standard_normal(m).astype(np.float32).astype(np.float64)
    present = np.zeros(N, dtype=bool)
    present[a] = True
    present[b] = True
    lone = np.nonzero(~present)[0]
    offsets = np.concatenate([2 * np.arange(m + 1), 2 * m + np.arange(1, len(lone) + 1)]).astype(np.int64)
    ids = np.concatenate([np.stack([a, b], axis=1).ravel(), lone]).astype(np.int32)
    coeffs = np.concatenate([J, np.zeros(len(lone))])
    return offsets, ids, coeffs


def random_pubo_arrays(n_vars, n_terms, fractions=(0.1, 0.4, 0.3, 0.2), cmax=10, seed=0):
    """random_pubo as a term CSR, vectorised: per term `deg` distinct ids
    drawn uniformly (rows with a repeated id are redrawn), integer
    coefficients uniform in [-cmax, cmax] \\ {0}, terms in random order."""
    rng = np.random.default_rng(seed)
    counts = [int(round(f * n_terms)) for f in fractions]
    counts[-1] = n_terms - sum(counts[:-1])
    rows = []
    for deg, cnt in enumerate(counts, start=1):
        x = rng.integers(0, n_vars, size=(cnt, deg))
        while deg > 1:
            s = np.sort(x, axis=1)
            bad = np.nonzero((s[:, 1:] == s[:, :-1]).any(axis=1))[0]
            if len(bad) == 0:
                break
            x[bad] = rng.integers(0, n_vars, size=(len(bad), deg))
        rows.append(x)
    vals = np.array([v for v in range(-cmax, cmax + 1) if v != 0])
    order = rng.permutation(n_terms)
    lens = np.concatenate([np.full(c, d) for d, c in enumerate(counts, start=1)])[order]
    flat = [r for x in rows for r in x]
    ids = np.concatenate([flat[k] for k in order]).astype(np.int32) if n_terms else np.zeros(0, np.int32)
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    coeffs = rng.choice(vals, size=n_terms).astype(np.float64)
    return offsets, ids, coeffs


def batched_er_instances(n=4096, seed=4096):
    """The uneven batch: `n` Erdos-Renyi instances with N log-uniform in
    [64, 4096], mean degree 6, J uniform in {-3..3} \\ {0} and one field term
    h in {-1, 0, 1} per spin.  Returns [(N, offsets, ids, coeffs), ...]."""
    sizes = log_uniform_sizes(n, 64, 4096, seed=seed)
    return [(int(N),) + er_ising_arrays(int(N), 6.0, 3, seed=seed * 16 + k) for k, N in enumerate(sizes)]


def request_stream(rate, duration, seed=5005, pa_extra=None, pa_share=0.5):
    """A serving stream: Poisson arrivals at `rate` per second over `duration`
    seconds.  70 % Ising ER (N log-uniform in [100, 50000], mean degree 8,
    J ~ N(0,1) fp32), 30 % random PUBO (V log-uniform in [500, 20000] vars,
    10 V terms of degree 1/2/3/4 at 10/40/30/20 %, c in [-10, 10] \\ {0}); each
    request is population annealing (population 512, beta 0 -> 3 over 200
    steps of 10 sweeps) with probability `pa_share`, else simulated annealing
    (64 restarts x 1000 sweeps, beta 0.1 -> 3.0).  Returns a list of
    {"t", "kind", "n", "target", "params", "arrays", "nominal_attempts"}."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        k = len(out)
        ising = rng.random() < 0.7
        pa = rng.random() < pa_share
        if ising:
            n = int(round(np.exp(rng.uniform(np.log(100), np.log(50000)))))
            arrays = er_ising_f32_arrays(n, 8.0, seed=seed * 1000 + k)
        else:
            n = int(round(np.exp(rng.uniform(np.log(500), np.log(20000)))))
            arrays = random_pubo_arrays(n, 10 * n, seed=seed * 1000 + k)
        out.append(dict(t=t, kind="ising" if ising else "pubo", n=n, arrays=arrays,
                        **stream_request(pa, n, 9000 + k, pa_extra)))
    return out


def stream_request(pa, n, seed, pa_extra=None):
    """Target, params and nominal spin-flip attempts of one request of the
    serving stream on `n` variables."""
    if pa:
        params = {"population": 512, "beta_start": 0.0, "beta_stop": 3.0, "sweeps": 200,
                  "sweeps_per_replica": 10, "seed": seed, **(pa_extra or {})}
        # nominal: the population fluctuates around 512
        return dict(target="populationannealing.cpu", params=params, nominal_attempts=512 * 200 * 10 * n)
    params = {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 64, "sweeps": 1000, "seed": seed}
    return dict(target="simulatedannealing.qiotoolkit", params=params, nominal_attempts=64 * 1000 * n)
