"""Path counts beyond the FP64 range: per-level scaling of sigma.

A chain of k diamonds (h_{i-1} - {a_i, b_i} - h_i) has 2^k shortest paths
from h_0 to h_k; at k = 1100 that is 2^1100, past FP64's 2^1024.  Wider
diamonds (w middle vertices) give w^k.  Every betweenness step must still
return Brandes' values: they are checked against exact Brandes computed here
with Python integers and fractions (path counts as big integers, dependencies
as rationals; Seq_Brandes/Brandes.c:224-372 in exact arithmetic, textbook and
per-predecessor formulas), farness bit-exact, BC within 1e-12 relative.
"""
from collections import deque
from fractions import Fraction

import numpy as np
import pytest

import bc_amd

pytestmark = pytest.mark.gpu


def diamond_chain(k, width=2):
    """Hubs 0..k, middle vertices after them; returns the CSR."""
    u, v = [], []
    nxt = k + 1
    for i in range(1, k + 1):
        for _ in range(width):
            u += [i - 1, nxt]
            v += [nxt, i]
            nxt += 1
    return bc_amd.CSR.from_edges(nxt, np.array(u, np.int32), np.array(v, np.int32))


def exact_brandes(g, sources, compat=False):
    rp, col, n = g.row_ptr, g.col, g.n
    bc = [Fraction(0)] * n
    far = []
    smax = 0
    for s in sources:
        dist = [-1] * n
        sigma = [0] * n
        npred = [0] * n
        dist[s], sigma[s] = 0, 1
        order, q = [], deque([s])
        while q:
            x = q.popleft()
            order.append(x)
            for y in col[rp[x]:rp[x + 1]]:
                if dist[y] < 0:
                    dist[y] = dist[x] + 1
                    q.append(y)
                if dist[y] == dist[x] + 1:
                    sigma[y] += sigma[x]
                    npred[y] += 1
        far.append(sum(d for d in dist if d > 0))
        smax = max(smax, max(sigma))
        delta = [Fraction(0)] * n
        for w in reversed(order):
            for v in col[rp[w]:rp[w + 1]]:
                if dist[v] == dist[w] - 1:
                    delta[v] += Fraction(sigma[v], sigma[w]) * (1 + delta[w])
            if w != s:
                bc[w] += delta[w] * (npred[w] if compat else 1)
    return np.array([float(x) for x in bc]), np.array(far, np.int64), smax


CASES = [("diamonds-1100", 1100, 2), ("triple-diamonds-700", 700, 3)]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def chain(request):
    name, k, w = request.param
    g = diamond_chain(k, w)
    rng = np.random.default_rng(k)
    src = np.concatenate([[0, k // 2, k], rng.choice(np.arange(1, g.n), 61, replace=False)]).astype(np.int32)
    want = {}
    for compat in (False, True):
        bc, far, smax = exact_brandes(g, src, compat)
        assert smax > 2 ** 1024  # the case is out of FP64 range unscaled
        want[compat] = (bc, far)
    print(f"[gpu] {name}: n={g.n}, exact Brandes done", flush=True)
    return name, g, src, want


@pytest.mark.parametrize("frontier", ["dense", "sparse", "source"])
@pytest.mark.parametrize("compat", [False, True])
def test_scaled_counts_match_exact_brandes(ctx, chain, frontier, compat):
    name, g, src, want = chain
    dg = ctx.upload(g)
    mode = bc_amd.CENT_BC_REF_COMPAT if compat else bc_amd.CENT_BC_TEXTBOOK
    print(f"[gpu] {name} {frontier} compat={compat}", flush=True)
    bc, far, _ = ctx.bc(dg, src, mode=mode, frontier=frontier)
    wbc, wfar = want[compat]
    assert np.array_equal(far, wfar)
    assert np.all(np.isfinite(bc))
    err = np.max(np.abs(bc - wbc)) / max(1.0, np.abs(wbc).max())
    assert err <= 1e-12, err


def test_scaled_counts_source_parallel_slot_reuse(ctx, chain):
    """Sources redone in scaled form share slots with unscaled ones."""
    name, g, src, want = chain
    dg = ctx.upload(g)
    bc, far, _ = ctx.bc(dg, src, frontier="source", max_slots=3)
    assert np.array_equal(far, want[False][1])
    assert np.max(np.abs(bc - want[False][0])) / max(1.0, np.abs(want[False][0]).max()) <= 1e-12


def test_scaled_all_sources_reduced_equals_plain(ctx):
    """All-source BC of a 1100-diamond chain: the reduced form (4-cycles,
    small counts) against the plain form (counts up to 2^1100, scaled)."""
    g = diamond_chain(1100, 2)
    dg = ctx.upload(g)
    bc, far, _ = ctx.bc(dg, None)
    red = bc_amd.Reduced(g)
    rbc, rfar = ctx.bc_reduced(red)
    assert np.array_equal(far, rfar)
    assert np.all(np.isfinite(bc))
    assert np.max(np.abs(bc - rbc)) / np.abs(rbc).max() <= 1e-12
    # hubs carry all pairs across them: BC(h_i) for ordered pairs
    # = 2 * (pairs separated by h_i) plus the paths through it inside diamonds
    assert rbc[550] > rbc[551 + 1100] > 0
