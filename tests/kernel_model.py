"""Literal per-thread model of the reference GPU-BSW kernels (test-only checker).

This is an independent restatement of the CUDA code, written from the kernel's
control flow rather than from a closed form: one CUDA block of ``T`` threads is
simulated diagonal by diagonal with its registers, the warp-boundary and
retired-lane spill arrays, the 32-lane ``__shfl_sync`` moves and the two-level
``__shfl_down_sync`` arg-max trees.  It is slow (pure Python) and only used on
tiny inputs to pin the oracle's tie rules.

Sources followed line by line:
  forward  kernel.cpp:198-394 (DNA) / 605-828 (protein)
  reverse  kernel.cpp:397-604 (DNA) / 830-1059 (protein)
  warp arg-max       kernel.cpp:47-86   (forward), 3-45  (reverse)
  block arg-max      kernel.cpp:137-182 (forward), 90-135 (reverse)
Assumption for partial warps (``T`` not a multiple of 32): a shuffle whose
source lane does not exist returns the caller's own value.
"""

WARP = 32


def _s16(x):
    """Assignment to ``short`` (two's complement wrap)."""
    x &= 0xFFFF
    return x - 0x10000 if x & 0x8000 else x


def _warp_reduce(vals, inds, inds2, nlanes, reverse):
    """kernel.cpp:47-86 (reverse=False) / 3-45 (reverse=True) for one warp.

    vals/inds/inds2 hold the 32 lanes' inputs (lanes >= nlanes do not exist).
    Returns lane 0's (val, ind, ind2)."""
    val = list(vals)
    my_max = list(vals)
    ind = list(inds)
    ind2 = list(inds2)
    off = WARP // 2
    while off > 0:
        old_val, old_ind, old_ind2 = list(val), list(ind), list(ind2)
        for lane in range(nlanes):
            src = lane + off
            if src >= WARP or src >= nlanes:
                src = lane  # out of range / absent lane: own value
            temp = old_val[src]
            v = max(old_val[lane], temp)
            n1, n2 = old_ind[src], old_ind2[src]
            val[lane] = v
            if v != my_max[lane]:
                ind[lane], ind2[lane], my_max[lane] = n1, n2, v
            elif v == temp:
                if (not reverse and n1 < ind[lane]) or (reverse and n2 > ind2[lane]):
                    ind[lane], ind2[lane] = n1, n2
        off //= 2
    return my_max[0], ind[0], ind2[0]


def _block_reduce(tmax, ti, tj, T, reverse):
    """kernel.cpp:137-182 / 90-135 for a block of T threads."""
    nwarps = (T + WARP - 1) // WARP
    loc = []
    for w in range(nwarps):
        lo = w * WARP
        n = min(WARP, T - lo)
        pad = WARP - n
        v = tmax[lo:lo + n] + [0] * pad
        a = ti[lo:lo + n] + [0] * pad
        b = tj[lo:lo + n] + [0] * pad
        loc.append(_warp_reduce(v, a, b, n, reverse))
    check = (32 + T - 1) // 32
    v, a, b = [], [], []
    n0 = min(WARP, T)
    for lane in range(WARP):
        if lane < check:
            v.append(loc[lane][0]); a.append(loc[lane][1]); b.append(loc[lane][2])
        else:
            v.append(0); a.append(-1); b.append(-1)
    return _warp_reduce(v, a, b, n0, reverse)


def _diag_sweep(longer, col_chars, max_size, min_size, T, score, open_, ext, reverse_rows):
    """The shared anti-diagonal loop (kernel.cpp:286-374 / 487-585).

    Returns per-thread (thread_max, thread_max_i, thread_max_j)."""
    valid_arr = [0] * min_size + [1] * min_size + [0] * min_size
    ptr = 2 * min_size
    cur = {"H": [0] * T, "E": [0] * T, "F": [0] * T}
    prv = {"H": [0] * T, "E": [0] * T, "F": [0] * T}
    pp = {"H": [0] * T, "E": [0] * T, "F": [0] * T}
    tmax, ti, tj = [0] * T, [0] * T, [0] * T
    irow = [1] * T
    nwarps = (T + WARP - 1) // WARP
    sh_E, sh_H, sh_PPH = [0] * nwarps, [0] * nwarps, [0] * nwarps
    sp_E, sp_H, sp_PPH = [0] * T, [0] * T, [0] * T
    for diag in range(max_size + min_size - 1):
        ptr -= 1 if (diag < min_size or diag >= max_size) else 0
        for k in ("H", "E", "F"):  # register rotation, 291-307
            tmp = prv[k]
            prv[k] = cur[k]
            cur[k] = pp[k]
            pp[k] = tmp
            cur[k] = [0] * T
        for t in range(T):
            if t % WARP == WARP - 1:
                w = t // WARP
                sh_E[w], sh_H[w], sh_PPH[w] = prv["E"][t], prv["H"][t], pp["H"][t]
            if diag >= max_size:
                sp_E[t], sp_H[t], sp_PPH[t] = prv["E"][t], prv["H"][t], pp["H"][t]
        newH, newE, newF = list(cur["H"]), list(cur["E"]), list(cur["F"])
        for t in range(min(T, min_size)):
            idx = ptr + t
            if not (0 <= idx < len(valid_arr) and valid_arr[idx]):
                continue
            lane, warp = t % WARP, t // WARP
            fval = _s16(prv["F"][t] + ext)
            hfval = _s16(prv["H"][t] + open_)
            src = t - 1 if lane > 0 else t - 1 + WARP  # shfl from laneId-1 (mod 32)
            e_sh = prv["E"][src] if 0 <= src < T else 0
            he_sh = prv["H"][src] if 0 <= src < T else 0
            if diag >= max_size:
                eval_ = _s16(sp_E[t - 1] + ext)
                heval = _s16(sp_H[t - 1] + open_)
            else:
                eval_ = _s16((sh_E[warp - 1] if (warp != 0 and lane == 0) else e_sh) + ext)
                heval = _s16((sh_H[warp - 1] if (warp != 0 and lane == 0) else he_sh) + open_)
            if warp == 0 and lane == 0:
                eval_, heval = 0, 0
            f = fval if fval > hfval else hfval
            e = eval_ if eval_ > heval else heval
            pph_sh = pp["H"][src] if 0 <= src < T else 0
            if diag >= max_size:
                fpph = sp_PPH[t - 1]
            else:
                fpph = sh_PPH[warp - 1] if (warp != 0 and lane == 0) else pph_sh
            if warp == 0 and lane == 0:
                fpph = 0
            i = irow[t]
            rchar = longer[max_size - i] if reverse_rows else longer[i - 1]
            d = _s16(fpph + score(rchar, col_chars[t]))
            h = max(d, f, e, 0)
            newH[t], newE[t], newF[t] = h, e, f
            if not reverse_rows:
                if not tmax[t] >= h:
                    ti[t], tj[t], tmax[t] = i, t + 1, h
            else:
                if not tmax[t] >= h:
                    ti[t], tj[t], tmax[t] = max_size - i, min_size - t - 1, h
            irow[t] = i + 1
        cur["H"], cur["E"], cur["F"] = newH, newE, newF
    return tmax, ti, tj


def forward(seqA, seqB, T, score, open_, ext):
    """Model of sequence_{dna,aa}_kernel for one block.  seqA = ref (contig),
    seqB = query (read).  Returns (score, ref_end, query_end)."""
    la, lb = len(seqA), len(seqB)
    max_size, min_size = max(la, lb), min(la, lb)
    if la < lb:
        cols, longer = seqA, seqB
    else:
        cols, longer = seqB, seqA
    col_chars = [cols[t] if t < min_size else None for t in range(T)]
    tmax, ti, tj = _diag_sweep(longer, col_chars, max_size, min_size, T, score, open_, ext, False)
    v, i, j = _block_reduce(tmax, ti, tj, T, reverse=False)
    if la < lb:
        return v, j, i  # seqA_end = j, seqB_end = i (381-386)
    return v, i, j


def reverse(seqA, seqB, endA, endB, T, score, open_, ext):
    """Model of sequence_{dna,aa}_reverse.  Returns (ref_begin, query_begin)."""
    max_size, min_size = max(endA, endB), min(endA, endB)
    if endA < endB:
        cols = [seqA[endA - 1 - t] if t < endA else None for t in range(T)]
        longer = seqB
    else:
        cols = [seqB[endB - 1 - t] if t < endB else None for t in range(T)]
        longer = seqA
    tmax, ti, tj = _diag_sweep(longer, cols, max_size, min_size, T, score, open_, ext, True)
    v, i, j = _block_reduce(tmax, ti, tj, T, reverse=True)
    if endA < endB:
        return j, i  # seqA_begin = j, seqB_begin = i (591-595)
    return i, j


def align(ref, query, score, open_, ext, T=None, T_rev=None):
    """Forward + reverse as the driver chains them (driver.cpp:176-210).
    T defaults to min(len) (the reference uses the batch maximum); T_rev to
    min(ref_end, query_end).  Returns (score, ref_begin, ref_end, query_begin,
    query_end)."""
    if T is None:
        T = max(1, min(len(ref), len(query)))
    s, ea, eb = forward(ref, query, T, score, open_, ext)
    if T_rev is None:
        T_rev = max(1, min(ea, eb))
    ba, bb = reverse(ref, query, ea, eb, T_rev, score, open_, ext)
    return s, ba, ea, bb, eb
