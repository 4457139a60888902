"""Python interpreter of the frame-form op stream -- TEST INFRASTRUCTURE ONLY.

Runs the plan that ``frame_plan(n, ops)`` compiles (the stream the GPU
kernel k_frame executes per 64-shot word group) on the CPU, word group by
word group, with coins and noise drawn by the oracle's counter-based draw,
so the host compiler can be checked against the oracle without a GPU.  The
LDS image is the kernel's: Rv [2n+1], the Pauli frame [2n], record words,
the all-ones word, then the noise windows.
"""
from __future__ import annotations

import numpy as np

import oracle as O

M64 = (1 << 64) - 1
OP_NOISE, OP_M = 10, 14
FLAG_CODES = (8, 9, 11, 12, 13, 15, 16)
FLAG_RANDOM = 0xFFFFFFFF


def noise_descs(ops):
    """(kind, t0, t1, t2) per noise number, in program order."""
    out = []
    for op in np.asarray(ops, np.int64).reshape(-1, 8):
        code, slot = int(op[0]), int(op[3]) & 0xFFFFFFFF
        if code == OP_NOISE:
            out.append(("pauli", int(op[4]), int(op[5]), int(op[6])))
        elif code in FLAG_CODES and slot == FLAG_RANDOM:
            out.append(("flag", int(op[4]), 0, 0))
    return out


def _masks(desc, seed, k, shots):
    """X-part and Z-part mask words of noise number k over 64 shots."""
    kind, t0, t1, t2 = desc
    mx = mz = 0
    for i, b in enumerate(shots):
        v = O.draw(seed, 2 * k + 1, b)
        if kind == "flag":
            bx, bz = (v >> 40) < t0, False
        else:
            u = v >> 32
            bx, bz = u < t1, t0 <= u < t2
        mx |= int(bx) << i
        mz |= int(bz) << i
    return mx, mz


def run(plan, n, B, seed, ops, meas_seq0=0, noise_seq0=0):
    """Returns (r [B][2n+1] int32, records {slot: (results[B], determined[B])})."""
    R = 2 * n + 1
    W = (B + 63) // 64
    descs = noise_descs(ops)
    stream = [int(v) for v in plan["stream"]]
    r_out = np.zeros((B, R), np.int32)
    rec_res, rec_det = {}, {}
    for w in range(W):
        shots = [64 * w + i for i in range(64)]
        nbufs = plan.get("nbufs", 1)
        size = plan["lds_noise"] + nbufs * plan["nb_words"]
        lds = [0] * size                      # Rv, the frame E (zero), records, ...
        lds[2 * n] = M64                      # scratch row r = 1 (init)
        lds[plan["lds_const"]] = M64           # the all-ones word (the zero word follows)
        if plan["preload_n"]:
            raise NotImplementedError("preloaded records")
        for ci, (cb, cw, k0, nk) in enumerate(plan["chunks"]):
            cb, cw, k0, nk = int(cb), int(cw), int(k0), int(nk)
            base = plan["lds_noise"] + (ci % nbufs) * plan["nb_words"]  # the chunk's window
            for j in range(nk):
                mx, mz = _masks(descs[k0 + j], seed, noise_seq0 + k0 + j, shots)
                lds[base + 2 * j] = mx
                lds[base + 2 * j + 1] = mz
            s = stream[cb:cb + cw]
            pos = 0
            while pos < cw:
                kind, h1 = s[pos], s[pos + 1]
                if kind == 4:
                    T, nbody = h1, s[pos + 2]
                    offs = s[pos + 8:pos + 8 + T + 1]
                    bstart = (pos + 8 + T + 2) & ~1
                    body = s[bstart:bstart + nbody]
                    assert 2 * offs[T] == nbody
                    written = set()
                    for t in range(T):
                        mine = set()
                        for e in range(offs[t], offs[t + 1]):
                            w0, w1 = body[2 * e], body[2 * e + 1]
                            dst = w0 & 0xFFFF
                            srcs = (w0 >> 16, w1 & 0xFFFF, w1 >> 16)
                            assert 2 * n + 1 <= dst < 2 * n + 1 + 2 * n
                            acc = 0
                            for v in srcs:
                                acc ^= lds[v]
                            lds[dst] ^= acc
                            mine.add(dst)
                        assert not (mine & written), "tasks of one layer share a word"
                        written |= mine
                    pos = bstart + nbody
                elif kind == 1:
                    T, ne = h1, s[pos + 2]
                    rows = s[pos + 8:pos + 8 + T]
                    offs = s[pos + 8 + T:pos + 8 + 2 * T + 1]
                    ents = s[pos + 8 + 2 * T + 1:pos + 8 + 2 * T + 1 + ne]
                    assert len(set(rows)) == T
                    for t in range(T):
                        acc = 0
                        for e in range(offs[t], offs[t + 1]):
                            acc ^= lds[ents[e]]
                        lds[rows[t]] ^= acc
                    pos += 8 + 2 * T + 1 + ne
                elif kind == 2:
                    p, slot, mi, Tc, ridx, ne = h1, s[pos + 2], s[pos + 3], s[pos + 4], s[pos + 5], s[pos + 6]
                    exq = s[pos + 7]
                    Rs = lds[n + p]
                    coin = 0
                    for i, b in enumerate(shots):
                        coin |= int((O.draw(seed, 2 * (meas_seq0 + mi), b) >> 40) < Tc) << i
                    for row in s[pos + 8:pos + 8 + ne]:
                        assert row not in (p, n + p) and row < 2 * n
                        lds[row] ^= Rs
                    lds[p] = Rs
                    lds[n + p] = coin ^ lds[exq]
                    rec_res.setdefault(slot, [0] * W)[w] = coin
                    rec_det.setdefault(slot, [0] * W)[w] = 0
                    if ridx != 0xFFFFFFFF:
                        lds[plan["lds_rec"] + ridx] = coin
                    pos += 8 + ne
                elif kind == 3:
                    G = h1
                    desc = s[pos + 8:pos + 8 + 4 * G + 1]
                    rows = s[pos + 8 + 4 * G + 1:]
                    for j in range(G):
                        slot, ridx, xq, o0 = desc[4 * j:4 * j + 4]
                        o1 = desc[4 * (j + 1) + 3] if j + 1 < G else desc[4 * G]
                        acc = 0
                        for e in range(o0, o1):
                            assert rows[e] != 2 * n
                            acc ^= lds[rows[e]]
                        res = acc ^ (M64 if xq >> 31 else 0) ^ lds[xq & 0xFFFF]
                        rec_res.setdefault(slot, [0] * W)[w] = res
                        rec_det.setdefault(slot, [0] * W)[w] = M64
                        if ridx != 0xFFFFFFFF:
                            lds[plan["lds_rec"] + ridx] = res
                        if (xq >> 30) & 1:
                            lds[2 * n] = acc
                    pos += 8 + 4 * G + 1 + desc[4 * G]
                else:
                    raise AssertionError(f"bad op kind {kind} at {cb + pos}")
        for b in range(64 * w, min(B, 64 * w + 64)):
            for row in range(R):
                r_out[b, row] = (lds[row] >> (b - 64 * w)) & 1
    recs = {}
    for slot in rec_res:
        res = np.array([(rec_res[slot][b // 64] >> (b % 64)) & 1 for b in range(B)], np.int32)
        det = np.array([(rec_det[slot][b // 64] >> (b % 64)) & 1 for b in range(B)], np.int32)
        recs[slot] = (res, det)
    return r_out, recs


def shared_xz(plan, n):
    """The final shared X/Z as int32 [2n+1][n] arrays."""
    R = 2 * n + 1
    x = np.zeros((R, n), np.int32)
    z = np.zeros((R, n), np.int32)
    for row in range(R):
        for q in range(n):
            x[row, q] = (int(plan["x"][row, q // 64]) >> (q % 64)) & 1
            z[row, q] = (int(plan["z"][row, q // 64]) >> (q % 64)) & 1
    return x, z
