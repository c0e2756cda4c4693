"""CPU emulation of the fused tile executor (damavand_amd/csrc/tile_ops.hpp),
driven by the exact plan the library would upload (dmv_circuit_plan_json).

Test infrastructure: it checks the scheduler's layouts, LDS swizzle offsets,
register renaming masks and store offsets on the CPU, index for index as the
kernels use them, so a wrong offset or mask is caught before any GPU run.
"""
import ctypes
import json

import numpy as np

from damavand_amd import _native as N

NO_REG = 0xFFFFFFFF
OP_GEN, OP_DIAG, OP_SWAP, OP_IPG, OP_IPZ = 0, 1, 2, 4, 6


def plan_json(handle, dtype=N.C128, adjoint=False, tile_qubits=0, tlog=0, max_gen=0):
    length = ctypes.c_uint64(0)
    N.call("dmv_circuit_plan_json", handle, dtype, int(adjoint), tile_qubits, tlog, max_gen, None, 0,
           ctypes.byref(length))
    buf = ctypes.create_string_buffer(length.value)
    N.call("dmv_circuit_plan_json", handle, dtype, int(adjoint), tile_qubits, tlog, max_gen, buf, length.value,
           ctypes.byref(length))
    return json.loads(buf.value.decode())


def swz(li, G):
    li = np.asarray(li, dtype=np.int64)
    f = np.zeros_like(li)
    s = G
    while s < 16:
        f ^= li >> s
        s += G
    return li ^ (f & ((1 << G) - 1))


def drop_bit(y, b):
    """An LDS byte offset with bit b removed, the bits above it shifted down."""
    y = np.asarray(y, dtype=np.int64)
    lo = (1 << b) - 1
    return (y & lo) | ((y >> 1) & ~lo)


def insert_zero(x, pos):
    low = x & ((1 << pos) - 1)
    return ((x >> pos) << (pos + 1)) | low


def highest_bit(m):
    for b in (16, 8, 4, 2, 1):
        if m >= b:
            return b
    return 1


def element_map(i, shift, blk, fold):
    """Element of amplitude i in the device layout, as include/damavand_gpu.h
    states it: with s = i >> shift, i ^ (P(s) << blk) ^ (F(s >> 6) << shift),
    P: bit j -> bit (7 j mod 11), F: bit j -> bit (6 - fold + j mod fold)."""
    i = np.asarray(i, dtype=np.int64)
    if not shift:
        return i
    span = i >> shift
    p = np.zeros_like(span)
    for j in range(16):
        p ^= ((span >> j) & 1) << ((7 * j) % 11)
    f = np.zeros_like(span)
    if fold:
        h = span >> 6
        for j in range(16):
            f ^= ((h >> j) & 1) << (6 - fold + j % fold)
    return i ^ (p << blk) ^ (f << shift)


class TileEmulator:
    def __init__(self, plan, elem=16):
        self.plan = plan
        self.n, self.k, self.rlog, self.tlog = plan["n"], plan["k"], plan["rlog"], plan["tlog"]
        self.NR, self.NT = 1 << self.rlog, 1 << self.tlog
        self.elem = elem
        self.G = 3 if elem == 16 else 4
        self.tids = np.arange(self.NT)

    def thread_local(self, lay):
        o = np.zeros(self.NT, dtype=np.int64)
        for j, b in enumerate(lay["thr"]):
            o |= ((self.tids >> j) & 1) << b
        return o

    def thread_global(self, gthr):
        o = np.zeros(self.NT, dtype=np.int64)
        for j, b in enumerate(gthr):
            o |= ((self.tids >> j) & 1) << b
        return o

    def tile_bases(self, P):
        t = np.arange(1 << (self.n - self.k), dtype=np.int64)
        for q in P["qpos"]:
            t = insert_zero(t, q)
        return t

    def transpose(self, v, A, src, B, dst, split=0):
        """v: [tiles, NT, NR]; writes register p at swz(lt_A)^src[p], reads r at swz(lt_B)^dst[r].
        split > 0: as the half-LDS kernels do it, two rounds (threads with top thread bit 0, then 1)
        through half the LDS, byte offsets with bit `split` removed (tables already without it)."""
        tiles = v.shape[0]
        la = swz(self.thread_local(A), self.G) * self.elem
        lb = swz(self.thread_local(B), self.G) * self.elem
        if split:
            la, lb = drop_bit(la, split), drop_bit(lb, split)
            rounds = [self.tids >> (self.tlog - 1) == h for h in (0, 1)]
            size = 1 << (self.k - 1)
        else:
            rounds = [np.ones(self.NT, dtype=bool)]
            size = 1 << self.k
        out = np.full_like(v, np.nan)
        for sel in rounds:
            lds = np.full((tiles, size), np.nan + 0j)  # nothing survives from the other round
            hits = np.zeros(size, dtype=np.int64)
            for p in range(self.NR):
                addr = la[sel] ^ src[p]
                assert np.all(addr % self.elem == 0)
                addr = addr // self.elem
                assert addr.max() < size, "transpose write beyond the kernel's LDS"
                lds[:, addr] = v[:, sel, p]
                np.add.at(hits, addr, 1)
            assert np.all(hits == 1), "transpose writes are not a bijection"
            for r in range(self.NR):
                addr = (lb[sel] ^ dst[r]) // self.elem
                assert addr.max() < size
                out[:, sel, r] = lds[:, addr]
        assert not np.isnan(out.real).any(), "transpose read an amplitude its round did not write"
        return out

    def apply(self, v, op, lam=None, sink=None):
        tm = (self.tids & op["ctrl_thr"]) == op["ctrl_thr"]
        kind = op["kind"]
        m = op["m"]
        m00, m01, m10, m11 = complex(m[0], m[1]), complex(m[2], m[3]), complex(m[4], m[5]), complex(m[6], m[7])
        en, fl, rm = op["en"], op["fl"], op["rmask"]
        if kind in (OP_IPG, OP_IPZ):
            # Im<lambda|A|psi>, A = ax X + ay Y + az Z with (ax, ay, az) = m[0..2]
            p_, l_ = v, lam
            sx = np.zeros(v.shape[:2])
            sy = np.zeros(v.shape[:2])
            sz = np.zeros(v.shape[:2])
            if kind == OP_IPZ:
                for r in range(self.NR):
                    if op["tbit"] == NO_REG:
                        sign = -1.0 if (rm >> r) & 1 else 1.0
                    else:
                        sign = np.where((self.tids >> op["tbit"]) & 1, -1.0, 1.0)
                    sz += sign * (l_[:, :, r].real * p_[:, :, r].imag - l_[:, :, r].imag * p_[:, :, r].real)
            else:
                hb = highest_bit(rm)
                for a in range(self.NR):
                    if a & hb:
                        continue
                    b = a ^ rm
                    # the member holding target bit 0 is a unless fl says otherwise
                    i0, i1 = (b, a) if (fl >> a) & 1 else (a, b)
                    u0, u1, v0, v1 = p_[:, :, i0], p_[:, :, i1], l_[:, :, i0], l_[:, :, i1]
                    sx += (np.conj(v0) * u1 + np.conj(v1) * u0).imag
                    sy += -(np.conj(v0) * u1).real + (np.conj(v1) * u0).real
                    sz += (np.conj(v0) * u0).imag - (np.conj(v1) * u1).imag
            acc = (m[0] * sx + m[1] * sy + m[2] * sz).sum(axis=1)
            sink(op["slot"], acc)
            return
        if kind == OP_DIAG:
            if op["plain"]:
                # the kernels flip signs instead of multiplying: valid for diag(1, -1) only
                assert list(m) == [1.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0], "plain flag on a general diagonal"
            for r in range(self.NR):
                if not (en >> r) & 1:
                    continue
                if op["tbit"] == NO_REG:
                    d = m11 if (rm >> r) & 1 else m00
                else:
                    d = np.where((self.tids >> op["tbit"]) & 1, m11, m00)[None, :]
                    d = np.broadcast_to(d, v.shape[:2])[:, tm]
                v[:, tm, r] = v[:, tm, r] * d
            return
        hb = highest_bit(rm)
        assert 0 < rm < self.NR
        if op["plain"]:
            # the kernel applies the matrix as is to every pair (p, p ^ rmask), p the first member;
            # the host claims that this is what the masks say
            first = sum(1 << q for q in range(self.NR) if not q & hb)
            assert kind == OP_GEN and en == (1 << self.NR) - 1 and (fl & first) == 0, "plain flag inconsistent"
            en, fl = (1 << self.NR) - 1, 0
        for p in range(self.NR):
            if p & hb:
                continue
            q = p ^ rm
            if not (en >> p) & 1:
                continue
            x, y = v[:, tm, p].copy(), v[:, tm, q].copy()
            if kind == OP_SWAP:
                v[:, tm, p], v[:, tm, q] = y, x
                continue
            a00, a01, a10, a11 = (m11, m10, m01, m00) if (fl >> p) & 1 else (m00, m01, m10, m11)
            v[:, tm, p] = a00 * x + a01 * y
            v[:, tm, q] = a10 * x + a11 * y

    def element(self, i):
        return element_map(i, *self.plan["layout"])

    def run_pass(self, P, psi, lam=None, sink=None):
        # psi / lam are device-layout buffers; addresses are formed as the kernels form them:
        # elem(tile base | thread offset) ^ greg[r], greg already mapped to elements by the planner
        g = self.tile_bases(P)
        gl = self.thread_global(P["gthr_ld"])
        idx_ld = self.element(g[:, None, None] | gl[None, :, None]) ^ np.array(P["greg_ld"], dtype=np.int64)[None, None, :]
        assert idx_ld.max() < psi.size
        v = psi[idx_ld].copy()
        lv = lam[idx_ld].copy() if lam is not None else None
        cur = P["lay_ld"]
        for ph in P["phases"]:
            if ph["transpose_in"]:
                v = self.transpose(v, cur, ph["src_lswz"], ph["lay"], ph["dst_lswz"], P["split"])
                if lv is not None:
                    lv = self.transpose(lv, cur, ph["src_lswz"], ph["lay"], ph["dst_lswz"], P["split"])
            cur = ph["lay"]
            for op in ph["ops"]:
                if op["kind"] in (OP_IPG, OP_IPZ):
                    self.apply(v, op, lam=lv, sink=sink)
                else:
                    self.apply(v, op)
                    if lv is not None:
                        self.apply(lv, op)
        if P["transpose_st"]:
            v = self.transpose(v, cur, P["st_src_lswz"], P["lay_st"], P["st_dst_lswz"], P["split"])
            if lv is not None:
                lv = self.transpose(lv, cur, P["st_src_lswz"], P["lay_st"], P["st_dst_lswz"], P["split"])
        gs = self.thread_global(P["gthr_st"])
        idx_st = self.element(g[:, None, None] | gs[None, :, None]) ^ np.array(P["greg_st"], dtype=np.int64)[None, None, :]
        flat = np.sort(idx_st.reshape(-1))
        assert np.array_equal(flat, np.arange(1 << self.n)), "stores do not cover the state exactly once"
        psi[idx_st] = v
        if lam is not None:
            lam[idx_st] = lv

    def to_device(self, amp):
        buf = np.full(amp.size, np.nan + 0j)
        buf[self.element(np.arange(amp.size))] = amp
        return buf

    def from_device(self, buf):
        return buf[self.element(np.arange(1 << self.n))]

    def forward(self, psi):
        buf = self.to_device(psi)
        for P in self.plan["passes"]:
            self.run_pass(P, buf)
        return self.from_device(buf)

    def reverse(self, psi, lam):
        bp, bl = self.to_device(psi), self.to_device(lam)
        grads = np.zeros(self.plan["n_slots"])

        def sink(slot, acc):
            grads[slot] += acc.sum()

        for P in self.plan["passes"]:
            self.run_pass(P, bp, bl, sink)
        return self.from_device(bp), self.from_device(bl), grads
