"""Test helpers: run the oracle's frame chain and compare outputs bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle


def to_oracle_uniforms(u) -> oracle.Uniforms:
    """Copy a product vrt_uniforms (same 304-byte layout) into the oracle's."""
    o = oracle.Uniforms()
    C.memmove(C.byref(o), C.byref(u), 304)
    return o


def f16(a: np.ndarray) -> np.ndarray:
    return np.frombuffer(np.ascontiguousarray(a, np.uint16).tobytes(), np.float16).reshape(a.shape).astype(np.float32)


class OracleChain:
    """The reference's frame loop (main.js:680-747) on the oracle: pathtrace ->
    temporal -> a-trous 1/2/4 with gbuf/accum ping-pong (parity = frame & 1)."""

    def __init__(self, grid: oracle.OracleGrid, w: int, h: int, fmt: int = 0, threads: int = 0):
        self.g, self.w, self.h, self.fmt, self.threads = grid, w, h, fmt, threads
        adt = np.uint16 if fmt == 0 else np.float32
        self.gbuf = [np.zeros((h, w, 4), np.float32) for _ in range(2)]
        self.accum = [np.zeros((h, w, 4), adt) for _ in range(2)]
        self.frames = 0

    def frame(self, u: oracle.Uniforms, fclamp: float, spp: int = 1):
        p = self.frames & 1
        rad, alb, gb, cnt = oracle.pathtrace(self.g, u, fclamp, spp, threads=self.threads)
        self.gbuf[p] = gb
        acc = oracle.temporal(u, rad, gb, self.accum[1 - p], self.gbuf[1 - p], self.fmt,
                              threads=self.threads)
        self.accum[p] = acc
        out = dict(radiance=rad, albedo=alb, gbuf=gb, accum=acc, counters=cnt)
        if u.params0[2] & 2:
            d0 = oracle.atrous(u, acc, gb, 1, self.fmt, threads=self.threads)
            d1 = oracle.atrous(u, d0, gb, 2, 0, threads=self.threads)
            d0 = oracle.atrous(u, d1, gb, 4, 0, threads=self.threads)
            out["denoised"] = d0
            out["hdr"] = oracle.hdr(d0, 0, alb)
            out["ldr"] = oracle.present(d0, 0, alb, u.params1[2])
        else:
            out["hdr"] = oracle.hdr(acc, self.fmt, alb)
            out["ldr"] = oracle.present(acc, self.fmt, alb, u.params1[2])
        self.frames += 1
        return out


def assert_bits_equal(name: str, got: np.ndarray, want: np.ndarray, max_report: int = 5):
    got = np.ascontiguousarray(got)
    want = np.ascontiguousarray(want)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    g = got.view(np.uint8).reshape(got.shape[:2] + (-1,)) if got.ndim == 3 else got.view(np.uint8)
    w = want.view(np.uint8).reshape(want.shape[:2] + (-1,)) if want.ndim == 3 else want.view(np.uint8)
    if np.array_equal(g, w):
        return
    bad = np.argwhere(np.any(g != w, axis=-1)) if got.ndim == 3 else np.argwhere(g != w)
    rows = "; ".join(f"{tuple(i)}: got {got[tuple(i)]} want {want[tuple(i)]}" for i in bad[:max_report])
    raise AssertionError(f"{name}: {len(bad)} elements differ bitwise, e.g. {rows}")
