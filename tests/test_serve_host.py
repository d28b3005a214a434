"""Host logic of the served form that needs no GPU: the HBM arena (first fit, 256-byte
blocks, merging of freed neighbours), the precision modes a Server accepts, and request
re-timing."""
import numpy as np
import pytest

from minifem.serve import Arena, Request, Server, retime


def test_arena_first_fit_and_merge():
    a = Arena(4096)
    b1 = a.alloc(100)   # rounded to 256
    b2 = a.alloc(1000)  # 1024
    b3 = a.alloc(256)
    assert b1 == (0, 256) and b2 == (256, 1024) and b3 == (1280, 256)
    assert a.used == 1536
    assert a.alloc(4096) is None
    a.free(b2)
    assert a.alloc(2000) == (1536, 2048)  # the hole of 1024 is too small: first fit after it
    a.free(b1)
    a.free(b3)  # b1, b2 and b3 merge back into one block at 0
    assert a._free[0] == (0, 1536)
    assert a.alloc(1536) == (0, 1536)


def test_arena_returns_to_empty():
    a = Arena(1 << 20)
    rng = np.random.default_rng(3)
    blocks = []
    for _ in range(200):
        if blocks and rng.random() < 0.45:
            a.free(blocks.pop(int(rng.integers(len(blocks)))))
        else:
            b = a.alloc(int(rng.integers(1, 20000)))
            if b is not None:
                blocks.append(b)
    for b in blocks:
        a.free(b)
    assert a.used == 0 and a._free == [(0, 1 << 20)]


def test_server_rejects_f32coef_precision():
    """MF_PREC_MIXED_F32COEF is picked per request; a server in that mode would read
    float64 coefficient fields as float."""
    with pytest.raises(ValueError):
        Server([], 1024, precision=2)


def test_retime_keeps_payloads():
    reqs = [Request(r, 0.001 * r, r % 3, "coef", np.full(4, float(r))) for r in range(50)]
    for r in reqs:
        r.host = object()
    out = retime(reqs, 1e5, seed=7)
    assert [q.rid for q in out] == [r.rid for r in reqs]
    assert all(q.host is r.host and q.mesh == r.mesh for q, r in zip(out, reqs))
    t = np.array([q.t_arrival for q in out])
    assert np.all(np.diff(t) > 0) and t[-1] < 0.01  # 50 arrivals at 100,000/s
