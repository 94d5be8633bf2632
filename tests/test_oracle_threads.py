"""The CPU oracle driven from several threads at once.

The CPU baseline runs one oracle instance per host core from a thread pool.  The
Gauss-Legendre table behind every order >= 2 call is process-wide and built on
first use, so the first calls of all threads meet in its construction, and a
table reset can land while other threads read it.  Each case runs in a fresh
interpreter, so the table starts unbuilt, and checks that every thread gets the
single-threaded result and that the process ends normally (a table freed twice
aborts it).
"""
from __future__ import annotations

import subprocess
import sys
import textwrap
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]

_SCRIPT = textwrap.dedent(
    """
    import sys, threading
    import numpy as np
    sys.path.insert(0, {oracle!r})
    import oracle as O

    THREADS, ROUNDS = 16, {rounds}
    rng = np.random.default_rng(7)
    S, D = {samples}, {dim}
    data = [rng.standard_normal((S, D, D)) * 0.5 for _ in range(THREADS)]
    cot = [rng.standard_normal((D, D)) for _ in range(THREADS)]
    got = [[None] * ROUNDS for _ in range(THREADS)]
    bar = threading.Barrier(THREADS)

    def work(t):
        for r in range(ROUNDS):
            bar.wait()
            if {reset} and t % 4 == 0:
                O.gl_reset()
            f = O.forward(8, data[t].copy(), 0.0, 1.0, op="sum", integrator="Boole")
            g = O.vjp(8, data[t].copy(), cot[t].copy(), 0.0, 1.0, op="sum", integrator="Boole") if {vjp} else None
            got[t][r] = (f, g)

    ts = [threading.Thread(target=work, args=(t,)) for t in range(THREADS)]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    for t in range(THREADS):
        f0 = O.forward(8, data[t].copy(), 0.0, 1.0, op="sum", integrator="Boole")
        g0 = O.vjp(8, data[t].copy(), cot[t].copy(), 0.0, 1.0, op="sum", integrator="Boole") if {vjp} else None
        for r in range(ROUNDS):
            f, g = got[t][r]
            assert np.array_equal(f, f0) and (g is None or np.array_equal(g, g0)), (t, r)
    print("threads ok")
    """
)


def _run(rounds: int, reset: bool, samples: int, dim: int, vjp: bool) -> None:
    code = _SCRIPT.format(oracle=str(ROOT / "oracle"), rounds=rounds, reset=reset, samples=samples, dim=dim,
                          vjp=vjp)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and "threads ok" in p.stdout, (p.returncode, p.stdout[-400:], p.stderr[-800:])


def test_first_use_from_many_threads():
    """All threads' first calls build the table at once (the CPU baseline's shape: one
    d = 8, S = 16385 instance per thread; an unguarded table aborted this every time)."""
    for _ in range(2):
        _run(rounds=1, reset=False, samples=16385, dim=8, vjp=False)


def test_table_reset_while_other_threads_read_it():
    """A quarter of the threads rebuild the table each round while the rest read it."""
    _run(rounds=25, reset=True, samples=1025, dim=3, vjp=True)
