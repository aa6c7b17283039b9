"""The batch kernel computes a proposal's duni draws in parallel, lane i
taking draw i of a ten-draw lookahead from one batch of table reads, with
the carry c after i+1 draws in closed form.  These tests restate that
method on the host and check it bit for bit against the oracle's sequential
duni (gcmc.f:2786-2799, itself pinned to the published RANMAR known answer
in test_oracle_math.py), over states that span every table wrap-around."""
import numpy as np
import pytest

Q = np.float32(16777216.0)


def state(d):
    st = d.st
    u = np.array(st.u[:], dtype=np.float32)
    return u, np.float32(st.c), np.float32(st.cd), np.float32(st.cm), int(st.ir), int(st.jr)


def lookahead(u, c, cd, cm, ir0, jr0, n=10):
    """Draw i (i < n) from the state before any of them, as lane i computes
    it: the table entries of draw i are those the sequential draws would
    read, and c after i+1 steps is (c - (i+1) cd) mod cm in units of 2^-24."""
    out = []
    icd, icm, ic = int(cd * Q), int(cm * Q), int(c * Q)
    for i in range(n):
        ir = ir0 - i
        if ir <= 0:
            ir += 97
        jr = jr0 - i
        if jr <= 0:
            jr += 97
        w = np.float32(u[ir - 1] - u[jr - 1])
        if w < 0:
            w = np.float32(w + np.float32(1.0))
        ci = (ic - (i + 1) * icd) % icm
        cc = np.float32(ci) * np.float32(1.0 / 16777216.0)
        d = np.float32(w - cc)
        if d < 0:
            d = np.float32(d + np.float32(1.0))
        out.append(float(d))
    return out


@pytest.mark.parametrize("seed", [(12, 34, 56, 78), (1, 2, 3, 4), (177, 177, 177, 1)])
def test_lookahead_equals_sequential_draws(oracle, seed):
    d = oracle.Duni(*seed)
    rng = np.random.default_rng(sum(seed))
    for _ in range(60):
        for _ in range(int(rng.integers(0, 40))):
            d()
        u, c, cd, cm, ir, jr = state(d)
        # the invariants the method needs
        assert (ir - jr) % 97 == 64
        assert np.all(u * Q == np.floor(u * Q)) and np.all((u >= 0) & (u < 1))
        assert c * Q == np.floor(c * Q) and 0 <= c < cm
        ahead = lookahead(u, c, cd, cm, ir, jr)
        seq = [d() for _ in range(10)]
        assert ahead == seq


def test_lookahead_needs_the_index_gap(oracle):
    """With the indices closer than ten draws (a state the generator never
    reaches) a batch would read an entry an earlier draw of the same batch
    overwrote; the invariant is what rules that out."""
    d = oracle.Duni(12, 34, 56, 78)
    for _ in range(5):
        d()
    u, c, cd, cm, ir, jr = state(d)
    # overwrite distance: draw i reads u(jr - i), draw j writes u(ir - j);
    # they meet for j - i = ir - jr (mod 97) = 64 -> only 33 draws back
    hits = [(i, j) for i in range(10) for j in range(i)
            if (jr - i - 1) % 97 == (ir - j - 1) % 97]
    assert hits == []
