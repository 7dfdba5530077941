"""An independent per-read model of the reference's read windows and BAQ tags.

Written for the oracle pins from the reference's CUDA/host source text, one
functor at a time, with its integer widths (uint16 windows, short HMM windows,
uint32 loop counters) emulated, and none of the oracle's code: it works on
plain Python lists of cigar events, where the oracle and the kernels use
cursors over packed ops.  It covers

  - cigar event expansion          device/cigar.cu:539-614
  - filter_if_cigar_malformed      device/read_filters.cu:157-258
  - read offsets, alignment window device/snp_filter.cu:44-159
  - filter_if_read_malformed       device/read_filters.cu:78-154
  - read window, adapter clipping  device/cigar.cu:98-344
  - soft-clip removal              device/cigar.cu:347-408
  - no-insertion window            device/cigar.cu:411-471
  - reference window               device/cigar.cu:473-534
  - error bits and num_errors      device/cigar.cu:616-737
  - HMM window and bandwidth       device/baq.cu:89-187
  - HMM setup (clamp, lengths)     device/baq.cu:260-354
  - flat BAQ, cap, recode          device/baq.cu:842-1081

The HMM itself is evaluated by an unbanded forward-backward written from the
model's definition (the caller passes it in).  Undefined reads of the reference
(a window search that runs off the read's events) are reported as None.
"""
import numpy as np

NONE16 = 0xFFFF
M_, I_, D_, S_ = "M", "I", "D", "S"
# cigar op code -> event class (cigar.cu:68-93)
EVCLASS = {0: M_, 7: M_, 8: M_, 1: I_, 3: I_, 2: D_, 5: D_, 6: D_, 4: S_}
OP = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4, "H": 5, "P": 6, "=": 7, "X": 8}

F_PAIRED, F_UNMAP, F_MATE_UNMAP, F_REVERSE, F_MATE_REVERSE = 1, 4, 8, 16, 32
F_SECONDARY, F_QCFAIL, F_DUP = 256, 512, 1024


class Undefined(Exception):
    """The reference reads outside the read's own events here."""


def read_length(ops):
    return sum(n for op, n in ops if EVCLASS[op] in (M_, I_, S_))


def expand(ops):
    """Per event: class, read coordinate, reference coordinate (NONE16 where absent)."""
    ev, rc, fc = [], [], []
    ro = fo = 0
    for op, n in ops:
        c = EVCLASS[op]
        for _ in range(n):
            ev.append(c)
            if c == M_:
                rc.append(ro); fc.append(fo); ro += 1; fo += 1
            elif c in (S_, I_):
                rc.append(ro); fc.append(NONE16); ro += 1
            else:
                rc.append(NONE16); fc.append(fo); fo += 1
    return ev, rc, fc


def cigar_ok(ops):
    """filter_if_cigar_malformed: no N, clips only at the ends (H outside S)."""
    state = 0   # 0 lead H, 1 lead S, 2 body, 3 trail S, 4 trail H
    for op, _ in ops:
        if op == OP["N"]:
            return False
        if op == OP["H"]:
            if state == 1:
                return False
            if state in (2, 3):
                state = 4
        elif op == OP["S"]:
            if state == 0:
                state = 1
            elif state == 2:
                state = 3
            elif state == 4:
                return False
        else:
            if state in (0, 1):
                state = 2
            elif state in (3, 4):
                return False
    return True


def last_offset(ops):
    """compute_read_offset_list + compute_alignment_window: reference offset of the last read base
    that has one (None when no base has)."""
    offs, off = [], 0
    for op, n in ops:
        if op in (0, 7, 8):
            for _ in range(n):
                offs.append(off); off += 1
        elif op in (1, 3):
            offs += [off] * n
        elif op in (2, 5, 6):
            off += n
        else:
            offs += [None] * n
    for o in reversed(offs):
        if o is not None:
            return o
    return None


def is_active(r, contig_len):
    """filter_invalid_reads + filter_malformed_reads for a read dict (ops, start, flags, mapq)."""
    if r["mapq"] in (0, 255):
        return False
    if r["flags"] & (F_DUP | F_QCFAIL | F_UNMAP | F_SECONDARY):
        return False
    if not cigar_ok(r["ops"]):
        return False
    L = read_length(r["ops"])
    lo = last_offset(r["ops"])
    if lo is None or lo == 0:          # window y <= x
        return False
    if r["start"] >= contig_len:
        return False
    if not r["ops"] or L == 0:
        return False
    return True


def _ref_window_for(ev, rc, fc, wx, wy):
    """The reference window search of compute_reference_window / get_current_reference_window."""
    n = len(ev)
    fx = fy = 0
    for i in range(n):
        if rc[i] == wx:
            while i < n and fc[i] == NONE16:
                i += 1
            if i == n:
                return NONE16, NONE16
            fx = fc[i]
            break
    found = False
    for i in range(n - 1, -1, -1):
        if rc[i] == wy:
            while fc[i] == NONE16 and i > 0:
                i -= 1
            if i == 0:
                return NONE16, NONE16
            fy = fc[i]
            found = True
            break
    if not found:
        raise Undefined("backward reference-window search ran past the read")
    return fx, fy


def _adapter(r, ev, rc, fc, x, y):
    """remove_adapters (cigar.cu:112-344): the clipped read window after adapter clipping."""
    flags, isize = r["flags"], r.get("isize", 0)
    if isize == 0 or not flags & F_PAIRED or flags & (F_UNMAP | F_MATE_UNMAP):
        return x, y
    rev, mrev = bool(flags & F_REVERSE), bool(flags & F_MATE_REVERSE)
    if rev == mrev:
        return x, y
    start, mstart = r["start"], r.get("mate_start", 0)
    if rev:
        if not r["stop"] > mstart:
            return x, y
        boundary = (mstart - 1) & 0xFFFFFFFF
    else:
        if not start <= mstart + isize:
            return x, y
        boundary = (start + abs(isize) + 1) & 0xFFFFFFFF
    fx, fy = _ref_window_for(ev, rc, fc, x, y)
    if boundary < start + fx or boundary > start + fy:
        return x, y
    n = len(ev)

    def read_coord(ref_coord, right):
        for e in range(n):
            if start + fc[e] == ref_coord:
                c = rc[e]
                if c == NONE16:
                    if right:
                        while e < n:
                            c = rc[e]
                            if c != NONE16:
                                break
                            e += 1
                    else:
                        while e > 0:
                            e -= 1
                            c = rc[e]
                            if c != NONE16:
                                break
                    if c == NONE16:
                        return -1
                return c
        return -1

    if rev:
        stop = read_coord(boundary, False)                 # hardClipByReferenceCoordinates(-1, boundary)
        x = max(x, (stop + 1) & 0xFFFF)                    # max<uint16>(x, adapter.y + 1)
    else:
        st = read_coord(boundary, True)
        y = min(y, (st - 1) & 0xFFFF)                      # min<uint16>(y, adapter.x - 1)
    return x, y


def windows(r, contig):
    """Read window, no-insertion window, reference window, error bits, num_errors, HMM window,
    bandwidth and whether BAQ drops the read, for an active read.  contig: list of 4-bit codes."""
    ops, L = r["ops"], read_length(r["ops"])
    ev, rc, fc = expand(ops)
    n = len(ev)
    x, y = 0, L - 1
    x, y = _adapter(r, ev, rc, fc, x, y)
    # remove_soft_clips
    ci, ro = 0, 0
    while ci < len(ops) and ro < L:
        op, ln = ops[ci]
        if op in (4, 5):
            if ro + ln > x:
                x = min((ro + ln) & 0xFFFF, y)
            ro += ln
            ci += 1
        else:
            break
    ci, ro = len(ops) - 1, L - 1
    while ci >= 0 and ro > x:
        op, ln = ops[ci]
        if op in (4, 5):
            d = (ro - ln) & 0xFFFFFFFF
            if d < y:
                y = max(d & 0xFFFF, x)
            ro = d
            ci -= 1
        else:
            break
    if ci < 0 and ro > x:
        raise Undefined("trailing clip loop ran past the read's ops")
    # no-insertion window (an unset side reads as 0)
    nix = niy = 0
    for e in range(n):
        if rc[e] == NONE16 or rc[e] < x or ev[e] == I_:
            continue
        nix = rc[e]
        break
    for e in range(n - 1, -1, -1):
        if rc[e] == NONE16 or rc[e] > y or ev[e] == I_:
            continue
        niy = rc[e]
        break
    fx, fy = _ref_window_for(ev, rc, fc, nix, niy)
    # error bits and num_errors (only the real op codes M, I, D are visited)
    snp, ins, dele = set(), set(), set()
    nerr = 0
    cur = 0
    e0 = 0
    ref = r["ref_at"]          # reference base at contig position p (N past the end)
    for op, ln in ops:
        if op == 0:
            for e in range(e0, e0 + ln):
                cur = rc[e]
                if ref(r["start"] + fc[e]) != r["seq"][cur]:
                    snp.add(cur)
                    if x <= cur <= y:
                        nerr += 1
        elif op == 1:
            cur = rc[e0]
            if x <= cur <= y:
                off = cur - 1 if not r["flags"] & F_REVERSE else cur + ln
                if 0 <= off <= y:
                    ins.add(off)
                    nerr += 1
        elif op == 2:
            cref = fc[e0]
            if fx <= cref <= fy:
                if not r["flags"] & F_REVERSE:
                    dele.add(cur)
                    nerr += 1
                else:
                    off = (cur + 1) & 0xFFFF
                    if off < L:
                        dele.add(off)
                        nerr += 1
        e0 += ln
    # HMM window (baq.cu:89-187)
    left_clip = x != 0
    right_clip = y != L - 1
    lins = 0 if left_clip else nix - x
    rins = 0 if right_clip else y - niy
    s16 = lambda v: ((v + 0x8000) & 0xFFFF) - 0x8000
    hx = s16(fx - s16(lins) - 3)
    hy = s16(fy + s16(rins) + 3)
    shift = -1 if (hx < 0 and r["start"] < abs(hx)) else 0
    needs_baq = nerr != 0
    dropped = needs_baq and ((r["start"] + hy + shift) & 0xFFFFFFFF) >= r["contig_len"]
    rl = hy - hx + 1
    ql = y - x + 1
    bw = max(rl, ql) & 0xFFFF                             # uint16 bandWidth
    if 7 < abs(rl - ql):
        bw = abs(rl - ql) + 3
    if bw > 7:
        bw = 7
    if bw < abs(rl - ql):
        bw = abs(rl - ql)
    return dict(x=x, y=y, nix=nix, niy=niy, fx=fx, fy=fy, snp=snp, ins=ins, dele=dele, nerr=nerr,
                hx=hx, hy=hy, bw=bw & 0xFFFF, dropped=dropped, ev=ev, rc=rc, fc=fc)


def map_quality(p):
    """(int)(-4.343 * log(1 - max) + .499), k > 100 -> 99, k < 4 -> 4 (baq.cu:673-674); None near a
    rounding boundary, where either integer is a correct reading of the real value."""
    import math
    if p >= 1.0:
        return 4
    real = -4.343 * math.log(1 - p) + 0.499
    if abs(real - round(real)) < 1e-6:
        return None
    k = int(real)
    return 99 if k > 100 else (4 if k < 4 else k)


def baq_tags(r, w, hmm):
    """BQ tags of one active read (list over its qualities; 0xFF where unset), or None when a MAP
    quality sits on a rounding boundary.  hmm(ref, query, qual, bw, blen) -> (states, posteriors)."""
    L = read_length(r["ops"])
    q = r["qual"]
    tags = [0xFF] * L
    if w["dropped"]:
        return tags
    x, y = w["x"], w["y"]
    if w["nerr"] == 0:
        for i in range(x, y + 1):
            tags[i] = 64                                   # read_flat_baq
        return tags
    hx, hy = w["hx"], w["hy"]
    rlen = hy - hx + 1
    clamp = hx < 0 and r["start"] < abs(hx)
    hoff = 0 if clamp else r["start"] + hx
    rlen_eff = rlen - 1 if clamp else rlen
    ref = [r["ref_at"](hoff + j) for j in range(rlen_eff)]
    query = [r["seq"][i] for i in range(x, y + 1)]
    qual = [q[i] for i in range(x, y + 1)]
    states, post = hmm(ref, query, qual, w["bw"], rlen)
    bq = {}
    for j, p in enumerate(post):
        mq = map_quality(p)
        if mq is None:
            return None
        bq[x + j] = mq
    st = {x + j: s for j, s in enumerate(states)}
    # cap_baq_qualities
    ev, rc = w["ev"], w["rc"]
    n = len(ev)
    if clamp:
        rshift, qshift = abs(hx), r["start"]
    else:
        rshift, qshift = 0, 0
    ref_off = hx - w["fx"] + rshift
    baq_start = 0
    for i in range(n):
        if rc[i] != NONE16 and rc[i] >= x:
            baq_start = i
            break

    def walk(update, out):
        readI = refI = numD = 0
        for i in range(baq_start, n):
            c = rc[i]
            if c != NONE16:
                if c < x:
                    continue
                if c > y:
                    break
            e = ev[i]
            if e == S_:
                refI += 1
            elif e == I_:
                if update:
                    out[c] = q[c]
                readI += 1
            elif e == D_:
                refI += 1
                numD += 1
            else:
                if update:
                    exp = (refI - ref_off + (i - numD - baq_start - readI) + qshift) & 0xFFFFFFFF
                    s = st[c] & 0xFFFFFFFF
                    if (s & 3) != 0 or (s >> 2) != exp:
                        out[c] = 4
                    else:
                        out[c] = min(bq[c], q[c])
                readI += 1
                refI += 1
        return readI

    val = dict(bq)
    if walk(False, None) == y - x + 1:
        walk(True, val)
    else:
        for i in range(x, y + 1):
            val[i] = q[i]
    for i in range(L):
        if i in val and val[i] != 0xFF:
            tags[i] = ((q[i] + 64) - val[i]) & 0xFF        # recode_baq_qualities
    return tags


# ---------------------------------------------------------------------------
# BAQ: an unbanded glocal HMM written from the model definition

EM, EI = 0.33333333333, 0.25


def hmm_independent(ref, query, qual, bw, blen=None):
    """Forward-backward over the full (query x reference) lattice with states
    M/I/D, the BAQ transition model, and cells outside |i - k| <= bw removed.
    Returns (MAP state, MAP posterior) per query base.  blen: the window length that sets the
    begin probabilities (the reference keeps the unclamped length there, baq.cu:283-286)."""
    lr, lq = len(ref), len(query)
    gop, gep = 10 ** -4.0, 0.1
    sM = 1.0 / (2 * lq + 2)
    bl = lr if blen is None else blen
    bM, bI = (1 - gop) / bl, gop / bl
    t = np.array([[(1 - 2 * gop) * (1 - sM), gop * (1 - sM), gop * (1 - sM)],
                  [(1 - gep) * (1 - sM), gep * (1 - sM), 0.0],
                  [1 - gep, 0.0, gep]])

    def emis(k, i):
        if ref[k - 1] == 15 or query[i - 1] == 15:
            return 1.0
        p = 10 ** (-max(qual[i - 1], 4) / 10.0)
        return 1 - p if ref[k - 1] == query[i - 1] else p * EM

    inband = lambda i, k: 1 <= k <= lr and abs(i - k) <= bw
    F = np.zeros((lq + 1, lr + 2, 3))
    s = np.zeros(lq + 2)
    for k in range(1, lr + 1):
        if inband(1, k):
            F[1, k, 0] = emis(k, 1) * bM
            F[1, k, 1] = EI * bI
    s[1] = F[1].sum()
    F[1] /= s[1]
    for i in range(2, lq + 1):
        for k in range(1, lr + 1):
            if not inband(i, k):
                continue
            F[i, k, 0] = emis(k, i) * (F[i - 1, k - 1] @ t[:, 0])
            F[i, k, 1] = EI * (F[i - 1, k, 0] * t[0, 1] + F[i - 1, k, 1] * t[1, 1])
            F[i, k, 2] = F[i, k - 1, 0] * t[0, 2] + F[i, k - 1, 2] * t[2, 2]
        s[i] = F[i].sum()
        F[i] /= s[i]
    s[lq + 1] = (F[lq, :, 0] * sM + F[lq, :, 1] * sM).sum()
    B = np.zeros((lq + 2, lr + 2, 3))
    for k in range(1, lr + 1):
        if inband(lq, k):
            B[lq, k, 0] = sM / s[lq] / s[lq + 1]
            B[lq, k, 1] = sM / s[lq] / s[lq + 1]
    for i in range(lq - 1, 0, -1):
        for k in range(lr, 0, -1):
            if not inband(i, k):
                continue
            e = emis(k + 1, i + 1) * B[i + 1, k + 1, 0] if k < lr else 0.0
            B[i, k, 0] = e * t[0, 0] + EI * t[0, 1] * B[i + 1, k, 1] + t[0, 2] * B[i, k + 1, 2]
            B[i, k, 1] = e * t[1, 0] + EI * t[1, 1] * B[i + 1, k, 1]
            B[i, k, 2] = (e * t[2, 0] + t[2, 2] * B[i, k + 1, 2]) * (1.0 if i > 1 else 0.0)
        B[i] /= s[i]
    states, post = [], []
    for i in range(1, lq + 1):
        z = F[i, :, :2] * B[i, :, :2]
        tot = z.sum()
        k, st = np.unravel_index(np.argmax(z), z.shape)
        states.append(int(k - 1) << 2 | int(st))
        post.append(z[k, st] / tot)
    return states, post
