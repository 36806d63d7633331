"""CPU oracle of the scoring path's host-side rows — TEST INFRASTRUCTURE ONLY.

The MSA parser is a pure-Python restatement (text handling only); the
floating-point functions are C (``host_oracle.c``, built into liboracle.so
with -ffp-contract=off) so their f32/f64 operations and libm calls match the
reference's.  Only ``tests/`` and ``bench.py``'s cpu_baseline leg may import
this module.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _p, aa2num, lib

GAP = 0xFF
_WS = " \t\n\v\f\r"
_SIGS = False


def _L():
    global _SIGS
    L = lib()
    if not _SIGS:
        P = C.c_void_p
        L.or_pssm_from_msa.restype = C.c_int
        L.or_pssm_from_msa.argtypes = [P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_float, C.c_float,
                                       P, P, P, P]
        _SIGS = True
    return L


def parse_msa(data: str):
    """parse_msa (src/profile.rs:372-449): (rows as a uint8 matrix, query_len).
    Raises ValueError on the reference's errors."""
    a2n = aa2num()
    raw, cur, started = [], [], False
    for line in data.split("\n"):
        line = line.rstrip(_WS)
        if line.startswith(">"):
            if started:
                raw.append("".join(cur))
                cur = []
            started = True
        else:
            cur.append(line.strip(_WS))
    if started:
        raw.append("".join(cur))
    if not raw:
        raise ValueError("MSA contains no sequences")
    is_a3m = any(any("a" <= ch <= "z" for ch in s) for s in raw)
    msa = []
    if is_a3m:
        for s in raw:
            row = []
            for ch in s:
                if "a" <= ch <= "z":
                    continue
                if ch in "-.":
                    row.append(GAP)
                elif "A" <= ch <= "Z":
                    row.append(int(a2n[ord(ch)]))
            msa.append(row)
    else:
        width = max(len(s) for s in raw)
        first = raw[0]
        cols = [i for i in range(width) if i < len(first) and first[i] not in "-."]
        for s in raw:
            row = []
            for i in cols:
                ch = s[i] if i < len(s) else "-"
                row.append(GAP if ch in "-." else int(a2n[ord(ch.upper()) & 0xFF]))
            msa.append(row)
    qlen = len(msa[0])
    for k, row in enumerate(msa):
        if len(row) != qlen:
            raise ValueError(f"MSA row {k + 1} has {len(row)} match columns, expected {qlen}")
    return np.array(msa, dtype=np.uint8).reshape(len(msa), qlen), qlen


def pssm_from_msa(msa: np.ndarray, qlen: int | None = None, pca=1.0, pcb=1.5, score_bias=0.0):
    """Profile::from_numeric_msa (src/profile.rs:90-182):
    (pssm int8 [L][20], prob f32 [L][20], neff f32 [L], consensus u8 [L])."""
    msa = np.ascontiguousarray(msa, dtype=np.uint8)
    n, stride = msa.shape
    L = stride if qlen is None else qlen
    pssm = np.zeros((max(L, 1), 20), np.int8)
    prob = np.zeros((max(L, 1), 20), np.float32)
    neff = np.zeros(max(L, 1), np.float32)
    cons = np.zeros(max(L, 1), np.uint8)
    rc = _L().or_pssm_from_msa(_p(msa), n, stride, L, pca, pcb, score_bias, _p(pssm), _p(prob), _p(neff), _p(cons))
    if rc != 0:
        raise ValueError("cannot build a profile from an empty MSA or with zero match columns")
    return pssm[:L], prob[:L], neff[:L], cons[:L]


def read_pssm_table(text: str):
    """Parse a PSSM text table (Profile::pssm_table layout,
    src/profile.rs:197-215): (int pssm [L][20], neff strings [L], query letters)."""
    lines = [l for l in text.split("\n") if l.strip()]
    header = lines[0].split()
    assert header[:2] == ["Pos", "Res"] and header[-1] == "Neff"
    rows, neff, res = [], [], []
    for l in lines[1:]:
        tok = l.split()
        res.append(tok[1])
        rows.append([int(x) for x in tok[2:22]])
        neff.append(tok[22])
    return np.array(rows, dtype=np.int32), neff, "".join(res)


# ---- E-value statistics (src/evalue.rs), f64 with the C library's exp/log ----

import math  # noqa: E402

# SEQUENCE_TABLE (src/evalue.rs:274-305): (open, ext, comp_bias) -> (lambda, K, alpha/lambda, beta)
SEQUENCE_TABLE = {(11, 1, False): (0.3666, 0.1177, 2.7, -13.0), (11, 1, True): (0.3649, 0.1076, 2.4, -12.0)}


def _fmax(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return a if a > b else b


def length_adjustment(k, adl, beta, m, n_res, n_seqs):
    """length_adjustment (src/evalue.rs:193-237)."""
    log_k = math.log(k)
    a = n_seqs
    mb = m * n_seqs + n_res
    c = n_res * m - _fmax(m, n_res) / k
    if c < 0.0:
        return 0.0
    ell_max = 2.0 * c / (mb + math.sqrt(_fmax(mb * mb - 4.0 * a * c, 0.0)))
    lo, hi, ell, conv = 0.0, ell_max, 0.0, False
    for i in range(1, 21):
        ss = (m - ell) * (n_res - n_seqs * ell)
        bar = adl * (log_k + math.log(_fmax(ss, 1.0))) + beta
        if bar >= ell:
            lo = ell
            if bar - lo <= 1.0:
                conv = True
                break
            if lo == hi:
                break
        else:
            hi = ell
        if lo <= bar <= hi:
            ell = bar
        elif i == 1:
            ell = hi
        else:
            ell = 0.5 * (lo + hi)
    out = float(math.floor(lo))
    if conv:
        up = float(math.ceil(lo))
        if up <= hi:
            ss = (m - up) * (n_res - n_seqs * up)
            if adl * (log_k + math.log(_fmax(ss, 1.0))) + beta >= up:
                out = up
    return _fmax(out, 0.0)


class EValueParams:
    """EValueParams (src/evalue.rs:311-461)."""

    def __init__(self, lam, k, db_residues, db_seqs, length=None):
        self.lam, self.k, self.log_k = lam, k, math.log(k)
        self.N = float(max(db_residues, 1))
        self.n = float(max(db_seqs, 1))
        self.length = length

    @staticmethod
    def sequence(open_, ext, comp_bias, db_residues, db_seqs):
        lam, k, adl, beta = SEQUENCE_TABLE[(open_, ext, bool(comp_bias))]
        return EValueParams(lam, k, db_residues, db_seqs, (adl, beta))

    def search_space(self, qlen):
        m = float(max(qlen, 1))
        if self.length is None:
            return m * self.N
        ell = length_adjustment(self.k, self.length[0], self.length[1], m, self.N, self.n)
        return _fmax(_fmax(m - ell, 1.0 / self.k), 1.0) * _fmax(_fmax(self.N - self.n * ell, self.N * 0.01), 1.0)

    def bit_score(self, raw):
        return (self.lam * float(raw) - self.log_k) / math.log(2.0)

    def evalue_in_space(self, raw, space):
        return self.k * space * math.exp(-self.lam * float(raw))

    def evalue(self, raw, qlen):
        return self.evalue_in_space(raw, self.search_space(qlen))

    def min_score_for(self, max_e, space):
        try:
            s = math.ceil(math.log(self.k * space / max_e) / self.lam)
        except (ValueError, OverflowError, ZeroDivisionError):
            return 1
        return int(min(max(float(s), 1.0), 2147483647 / 2.0))


def gumbel_mu_given_lambda(xs, lam):
    xmin = min(xs)
    s = 0.0
    for x in xs:
        s += math.exp(-lam * (x - xmin))
    return xmin - math.log(s / len(xs)) / lam


def fit_gumbel_ml(xs):
    """fit_gumbel_ml (src/evalue.rs:139-182): (lambda, mu) or None."""
    n = len(xs)
    if n < 10:
        return None
    mean = 0.0
    for x in xs:
        mean += x
    mean /= n
    var = 0.0
    for x in xs:
        var += (x - mean) * (x - mean)
    var /= (n - 1.0)
    if var != var or var <= 0.0:
        return None
    xmin = min(xs)

    def g(lam):
        s0 = s1 = 0.0
        for x in xs:
            w = math.exp(-lam * (x - xmin))
            s0 += w
            s1 += x * w
        return 1.0 / lam - mean + s1 / s0

    lam0 = math.pi / (math.sqrt(var) * math.sqrt(6.0))
    lo, hi = lam0 / 10.0, lam0 * 10.0
    if g(lo) < 0.0 or g(hi) > 0.0:
        return None
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) > 0.0:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-12 * hi:
            break
    lam = 0.5 * (lo + hi)
    return lam, gumbel_mu_given_lambda(xs, lam)


def rust_e2(x: float) -> str:
    """Rust's {:.2e}: C's %.2e mantissa (correctly rounded, ties to even),
    exponent without '+' or zero padding (src/search.rs:203-216)."""
    s = "%.2e" % x
    if "e" not in s:
        return {"nan": "NaN", "inf": "inf", "-inf": "-inf"}.get(s, s)
    mant, e = s.split("e")
    return f"{mant}e{int(e)}"


# ---- composition bias and low-complexity mask (pure Python, f64) -------------

def _round_half_away(x: float) -> int:
    return int(math.floor(x + 0.5)) if x >= 0.0 else -int(math.floor(-x + 0.5))


def local_bias_correction(q, penalize_only=False):
    """local_bias_correction_mode (src/compbias.rs:38-70)."""
    from . import blosum62
    sc, pb, _, _ = blosum62()
    sc = np.asarray(sc, dtype=np.int64).reshape(21, 21)
    exp_ = []
    for a in range(21):
        e = 0.0
        for b in range(20):
            e += float(pb[b]) * float(sc[a, b])
        exp_.append(e)
    n = len(q)
    out = []
    for i in range(n):
        lo, hi = max(i - 20, 0), min(i + 20, n)
        a = int(q[i])
        s = sum(int(sc[a, int(q[j])]) for j in range(lo, hi) if j != i)
        d = _round_half_away(exp_[min(a, 20)] - s / float(hi - lo))
        out.append(min(d, 0) if penalize_only else d)
    return np.array(out, dtype=np.int32)


def low_complexity_mask(seq, window=12, locut=2.2, hicut=2.5):
    """low_complexity_mask_into (src/mask.rs:70-134)."""
    n, w = len(seq), window
    mask = np.zeros(n, bool)
    if w < 2 or n < w:
        return mask
    clog = [0.0] + [c * math.log2(c) for c in range(1, w + 1)]
    log_w = math.log2(w)
    nwin = n - w + 1
    ent = [0.0] * nwin
    counts = [0] * 32
    s = 0.0
    sym = lambda r: min(int(r), 31)  # noqa: E731
    for r in seq[:w]:
        k = sym(r)
        s -= clog[counts[k]]
        counts[k] += 1
        s += clog[counts[k]]
    ent[0] = log_w - s / w
    for i in range(1, nwin):
        o = sym(seq[i - 1])
        s -= clog[counts[o]]
        counts[o] -= 1
        s += clog[counts[o]]
        k = sym(seq[i + w - 1])
        s -= clog[counts[k]]
        counts[k] += 1
        s += clog[counts[k]]
        ent[i] = log_w - s / w
    i = 0
    while i < nwin:
        if ent[i] <= locut + 1e-12:
            lo = i
            while lo > 0 and ent[lo - 1] <= hicut + 1e-12:
                lo -= 1
            hi = i
            while hi + 1 < nwin and ent[hi + 1] <= hicut + 1e-12:
                hi += 1
            mask[lo:hi + w] = True
            i = hi + 1
        else:
            i += 1
    return mask


def query_rows(q, comp_bias=True, penalize_only=False, mask=(15, 2.3, 2.3)):
    """search_query's rows (src/search.rs:432-463)."""
    from . import blosum62, seq_rows
    rows = np.asarray(seq_rows(q), dtype=np.int32).copy()
    if comp_bias:
        rows += local_bias_correction(q, penalize_only)[:, None]
    if mask is not None:
        mk = low_complexity_mask(q, *mask)
        sc = np.asarray(blosum62()[0], dtype=np.int32).reshape(21, 21)
        rows[mk] = sc[20]
    return rows


# ---- calibrate_profile's random targets (src/evalue.rs:52-114, 714-749) -------

def stream_seed(seed, c, i):
    from . import SplitMix64
    M = 0xFFFFFFFFFFFFFFFF
    return SplitMix64((seed ^ ((c * 0xA24BAED4963EE407) & M) ^ ((i * 0x9FB21C651E98DF25) & M)) & M).next_u64()


def calibration_targets(lengths, samples, seed):
    """The reference's per-sample targets: list of uint8 arrays."""
    from . import _p as P, lib as L
    out = []
    for i in range(samples):
        st = C.c_uint64(stream_seed(seed, 0xBEEF, i))
        n = max(int(lengths[L().or_splitmix_next(C.byref(st)) % len(lengths)]), 1)
        t = np.zeros(n, np.uint8)
        L().or_background_sequence(C.byref(st), n, P(t))
        out.append(t)
    return out


def calibrate_profile(rows, open_, ext, lengths, samples, seed):
    """(lambda, K, (lambda, mu)) with the oracle's exact scores and fit."""
    from . import gotoh_score
    xs = [float(gotoh_score(rows, t, open_, ext)) for t in calibration_targets(lengths, samples, seed)]
    f = fit_gumbel_ml(xs)
    if f is None:
        return None
    lam, mu = f
    mean_len = float(sum(int(x) for x in lengths)) / len(lengths)
    return lam, math.exp(lam * (mu + 0.5)) / (len(rows) * mean_len), (lam, mu)


# ---- calibrate_sequence's simulation and fit (src/evalue.rs:537-700) ----------

def simulate_pairs(seed, cell_id, m, n, samples):
    """simulate_scores' pairs: sample i's query (m) then target (n) from
    SplitMix64(stream_seed(seed, cell_id, i)); two uint8 arrays
    [samples, m] and [samples, n]."""
    from . import _p as P, lib as L
    qs = np.zeros((samples, m), np.uint8)
    ts = np.zeros((samples, n), np.uint8)
    for i in range(samples):
        st = C.c_uint64(stream_seed(seed, cell_id, i))
        L().or_background_sequence(C.byref(st), m, P(qs[i]))
        L().or_background_sequence(C.byref(st), n, P(ts[i]))
    return qs, ts


def local_bias_correction_rows(qs):
    """local_bias_correction (src/compbias.rs:38-70) of many queries of one
    length at once ([samples, m] codes -> [samples, m] int32): the window
    sums are exact integers, then the same f64 steps per position."""
    from . import blosum62
    sc, pb, _, _ = blosum62()
    sc = np.asarray(sc, dtype=np.int64).reshape(21, 21)
    exp_ = np.zeros(21)
    for a in range(21):
        e = 0.0
        for b in range(20):
            e += float(pb[b]) * float(sc[a, b])
        exp_[a] = e
    qs = np.asarray(qs, dtype=np.int64)
    s, n = qs.shape
    tot = np.zeros((s, n), np.int64)
    for d in range(-20, 20):
        if d == 0:
            continue
        lo, hi = max(0, -d), min(n, n - d)
        if lo < hi:
            tot[:, lo:hi] += sc[qs[:, lo:hi], qs[:, lo + d:hi + d]]
    i = np.arange(n)
    wl = (np.minimum(i + 20, n) - np.maximum(i - 20, 0)).astype(np.float64)
    delta = exp_[np.minimum(qs, 20)] - tot.astype(np.float64) / wl[None, :]
    out = np.where(delta >= 0.0, np.floor(delta + 0.5), -np.floor(-delta + 0.5))
    return out.astype(np.int32)


def fit_calibration(cells, lmin=200):
    """fit_calibration over [(m, n, scores)]: ((lambda, K, alpha/lambda,
    beta), cells [k, 8], ln K rms) or None."""
    from . import _p as P, lib as L
    cells = sorted(cells, key=lambda c: (c[0], c[1]))
    nc = len(cells)
    cm = np.array([c[0] for c in cells], np.uint64)
    cn = np.array([c[1] for c in cells], np.uint64)
    off = np.zeros(nc + 1, np.uint64)
    off[1:] = np.cumsum([len(c[2]) for c in cells])
    xs = np.ascontiguousarray(np.concatenate([np.asarray(c[2], np.float64) for c in cells]), dtype=np.float64)
    out6 = np.zeros(6)
    co = np.zeros((max(nc, 1), 8))
    f = L().or_fit_calibration
    f.argtypes = [C.c_void_p] * 4 + [C.c_size_t, C.c_uint64, C.c_void_p, C.c_void_p]
    if f(P(cm), P(cn), P(off), P(xs), nc, lmin, P(out6), P(co)) != 0:
        return None
    return (out6[0], out6[1], out6[2], out6[3]), co[:nc], out6[4]
