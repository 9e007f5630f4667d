"""Seeded synthetic workloads for the five benchmark configurations.

Every generator returns a ``Workload``: packed ref and query bytes with n+1
uint64 offsets (pair k = ref k against query k), the alphabet and the scoring.
Choices the configuration text leaves open are fixed here:
  * lognormal sigma = 0.5 for every lognormal length;
  * protein query substring length = U[0.5, 1.0] x ref length (clipped 50-1024);
  * indels are half single-residue deletions, half single-residue insertions of
    a background-drawn residue after the position;
  * BLOSUM62-conditional substitution p(b|a) proportional to p_b * 2^(s(a,b)/2),
    b != a, with p the background frequencies below;
  * seeds: 0x5EED0000 + config index unless given.
Background amino-acid frequencies: the BLOSUM62 target-frequency marginals
(Henikoff & Henikoff 1992) as tabulated by NCBI, rounded to 3 digits.
"""
from dataclasses import dataclass

import numpy as np

from .scoring import AA_ORDER, BLOSUM62, DNA_WORKLOAD_SCORES, AA_GAP_OPEN, AA_GAP_EXTEND

DNA_BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
AA20 = AA_ORDER[:20]
AA20_BYTES = np.frombuffer(AA20.encode(), dtype=np.uint8)
AA_BACKGROUND = np.array([0.074, 0.052, 0.045, 0.054, 0.025, 0.034, 0.054, 0.074, 0.026, 0.068,
                          0.099, 0.058, 0.025, 0.047, 0.039, 0.057, 0.051, 0.013, 0.032, 0.073])
AA_BACKGROUND = AA_BACKGROUND / AA_BACKGROUND.sum()

SEED_BASE = 0x5EED0000


@dataclass
class Workload:
    name: str
    alphabet: int          # 0 DNA, 1 protein
    ref: np.ndarray        # uint8
    ref_off: np.ndarray    # uint64, n+1
    query: np.ndarray      # uint8
    query_off: np.ndarray  # uint64, n+1
    scoring: tuple         # DNA: (match, mismatch, open, extend); protein: (open, extend)

    @property
    def n(self):
        return len(self.ref_off) - 1

    def lengths(self):
        return np.diff(self.ref_off).astype(np.int64), np.diff(self.query_off).astype(np.int64)

    def cells(self):
        lr, lq = self.lengths()
        return int((lr * lq).sum())

    def subset(self, idx):
        """Workload with pairs idx (in that order)."""
        idx = np.asarray(idx, dtype=np.int64)
        rb, ro = _gather(self.ref, self.ref_off, idx)
        qb, qo = _gather(self.query, self.query_off, idx)
        return Workload(self.name, self.alphabet, rb, ro, qb, qo, self.scoring)

    def pair(self, k):
        r = bytes(self.ref[self.ref_off[k]:self.ref_off[k + 1]])
        q = bytes(self.query[self.query_off[k]:self.query_off[k + 1]])
        return r, q


def _offsets(lengths):
    off = np.zeros(len(lengths) + 1, dtype=np.uint64)
    np.cumsum(lengths, out=off[1:])
    return off


def _gather(buf, off, idx):
    lens = (off[idx + 1] - off[idx]).astype(np.int64)
    new_off = _offsets(lens)
    total = int(new_off[-1])
    if total == 0:
        return np.zeros(1, dtype=np.uint8), new_off
    starts = np.repeat(off[idx].astype(np.int64), lens)
    within = np.arange(total, dtype=np.int64) - np.repeat(new_off[:-1].astype(np.int64), lens)
    return buf[starts + within], new_off


def _lognormal_lengths(rng, n, median, lo, hi, sigma=0.5):
    x = rng.lognormal(mean=np.log(median), sigma=sigma, size=n)
    return np.clip(np.rint(x), lo, hi).astype(np.int64)


def _substrings(rng, ref, ref_off, qlen):
    """Random substring of each ref with the requested lengths (<= ref length)."""
    rlen = np.diff(ref_off).astype(np.int64)
    qlen = np.minimum(qlen, rlen)
    start = (rng.random(len(rlen)) * (rlen - qlen + 1)).astype(np.int64)
    q_off = _offsets(qlen)
    total = int(q_off[-1])
    base = np.repeat(ref_off[:-1].astype(np.int64) + start, qlen)
    within = np.arange(total, dtype=np.int64) - np.repeat(q_off[:-1].astype(np.int64), qlen)
    return ref[base + within], q_off


def _mutate(rng, buf, off, sub_rate, indel_rate, substitute, draw_background):
    """Substitutions at sub_rate, then single-residue indels at indel_rate
    (half deletions, half insertions after the residue)."""
    n = len(off) - 1
    lens = np.diff(off).astype(np.int64)
    pair = np.repeat(np.arange(n, dtype=np.int32), lens)
    out = buf.copy()
    sub = rng.random(len(out), dtype=np.float32) < sub_rate
    out[sub] = substitute(rng, out[sub])
    r = rng.random(len(out), dtype=np.float32)
    dele = r < indel_rate / 2
    ins = (r >= indel_rate / 2) & (r < indel_rate)
    reps = (~dele).astype(np.int8) + ins.astype(np.int8)
    new = np.repeat(out, reps)
    new_pair = np.repeat(pair, reps)
    # the second copy of an inserted position becomes a random residue
    first = np.cumsum(reps, dtype=np.int64) - reps  # index in `new` of each source position's first copy
    ins_slots = first[ins] + 1
    new[ins_slots] = draw_background(rng, len(ins_slots))
    new_lens = np.bincount(new_pair, minlength=n).astype(np.int64)
    return new, _offsets(new_lens)


def _dna_substitute(rng, old):
    code = np.searchsorted(DNA_BASES, old).astype(np.uint8)
    return DNA_BASES[(code + rng.integers(1, 4, size=len(old), dtype=np.uint8)) % 4]


def _dna_background(rng, k):
    return DNA_BASES[rng.integers(0, 4, size=k, dtype=np.uint8)]


_AA_COND = None


def _aa_conditional():
    global _AA_COND
    if _AA_COND is None:
        cond = np.zeros((20, 20))
        for a in range(20):
            for b in range(20):
                if a != b:
                    cond[a, b] = AA_BACKGROUND[b] * 2.0 ** (BLOSUM62[a * 24 + b] / 2.0)
            cond[a] /= cond[a].sum()
        _AA_COND = np.cumsum(cond, axis=1)
    return _AA_COND


def _aa_substitute(rng, old):
    cum = _aa_conditional()
    code = np.full(len(old), -1, dtype=np.int64)
    lut = np.full(256, -1, dtype=np.int64)
    lut[AA20_BYTES] = np.arange(20)
    code = lut[old]
    u = rng.random(len(old))
    new = np.empty(len(old), dtype=np.uint8)
    for a in range(20):
        m = code == a
        if m.any():
            new[m] = AA20_BYTES[np.minimum(np.searchsorted(cum[a], u[m], side="right"), 19)]
    return new


def _aa_background(rng, k):
    return AA20_BYTES[rng.choice(20, size=k, p=AA_BACKGROUND)]


def _random_dna(rng, lens):
    off = _offsets(lens)
    return DNA_BASES[rng.integers(0, 4, size=max(int(off[-1]), 1), dtype=np.uint8)], off


def _random_protein(rng, lens):
    off = _offsets(lens)
    return AA20_BYTES[rng.choice(20, size=max(int(off[-1]), 1), p=AA_BACKGROUND)], off


def plain_dna(n=10_000, seed=None):
    """Config 0: ref ACGT U[200,1000]; query = substring U[100,150], 10% subs, 2% indels."""
    rng = np.random.default_rng(SEED_BASE + 0 if seed is None else seed)
    rlen = rng.integers(200, 1001, size=n)
    ref, ro = _random_dna(rng, rlen)
    qlen = rng.integers(100, 151, size=n)
    sub, so = _substrings(rng, ref, ro, qlen)
    q, qo = _mutate(rng, sub, so, 0.10, 0.02, _dna_substitute, _dna_background)
    return Workload("plain_dna_10k", 0, ref, ro, q, qo, DNA_WORKLOAD_SCORES)


def plain_protein(n=20_000, seed=None):
    """Config 1: background residues, ref lognormal(350) clip [50,1024]; query =
    substring (U[0.5,1] x ref) with 30% BLOSUM62-conditional substitutions and 3% indels."""
    rng = np.random.default_rng(SEED_BASE + 1 if seed is None else seed)
    rlen = _lognormal_lengths(rng, n, 350, 50, 1024)
    ref, ro = _random_protein(rng, rlen)
    qlen = np.clip(np.rint(rlen * rng.uniform(0.5, 1.0, size=n)), 50, 1024).astype(np.int64)
    sub, so = _substrings(rng, ref, ro, qlen)
    q, qo = _mutate(rng, sub, so, 0.30, 0.03, _aa_substitute, _aa_background)
    return Workload("plain_protein_20k", 1, ref, ro, q, qo, (AA_GAP_OPEN, AA_GAP_EXTEND))


def batched_dna(n=1_000_000, seed=None):
    """Config 2: query lognormal(150) clip [32,1024], ref lognormal(600) clip
    [64,2048], query <= ref, config-0 mutation model."""
    rng = np.random.default_rng(SEED_BASE + 2 if seed is None else seed)
    qlen = _lognormal_lengths(rng, n, 150, 32, 1024)
    rlen = _lognormal_lengths(rng, n, 600, 64, 2048)
    rlen = np.maximum(rlen, qlen)
    ref, ro = _random_dna(rng, rlen)
    sub, so = _substrings(rng, ref, ro, qlen)
    q, qo = _mutate(rng, sub, so, 0.10, 0.02, _dna_substitute, _dna_background)
    return Workload("batched_dna_1m", 0, ref, ro, q, qo, DNA_WORKLOAD_SCORES)


def long_protein(n=50_000, seed=None):
    """Config 3: query U[1024,4096], ref U[4096,16384], config-1 residue and mutation model."""
    rng = np.random.default_rng(SEED_BASE + 3 if seed is None else seed)
    rlen = rng.integers(4096, 16385, size=n)
    ref, ro = _random_protein(rng, rlen)
    qlen = rng.integers(1024, 4097, size=n)
    sub, so = _substrings(rng, ref, ro, qlen)
    q, qo = _mutate(rng, sub, so, 0.30, 0.03, _aa_substitute, _aa_background)
    return Workload("long_protein_50k", 1, ref, ro, q, qo, (AA_GAP_OPEN, AA_GAP_EXTEND))


CONFIGS = {0: plain_dna, 1: plain_protein, 2: batched_dna, 3: long_protein}


def serving_requests(n_requests, seed=None, dna_fraction=0.7, max_pairs=4096):
    """Config 4 request stream: sizes log-uniform in [1, max_pairs]; each request
    is DNA (config-2 generator) with probability dna_fraction, else protein
    (config-1 generator).  Returns a list of Workloads (arrival times are drawn by
    the load generator)."""
    rng = np.random.default_rng(SEED_BASE + 4 if seed is None else seed)
    reqs = []
    for r in range(n_requests):
        size = int(np.exp(rng.uniform(0.0, np.log(max_pairs))))
        sub_seed = int(rng.integers(0, 2**31))
        if rng.random() < dna_fraction:
            w = batched_dna(size, seed=sub_seed)
        else:
            w = plain_protein(size, seed=sub_seed)
        w.name = "req%d" % r
        reqs.append(w)
    return reqs
