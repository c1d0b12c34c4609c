"""Checkpoint schedules for the reverse sweep of an unsteady primal run.

The reverse sweep needs the primal state u_n at the start of every step n,
in reverse order.  Three schedules trade HBM for forward recomputation:

  store_all   keep every u_n (and, optionally, every step tape): no
              recomputation; nSteps states.
  uniform     keep u at every `interval`-th step, re-run each segment
              forward before reversing it: the reference's schedule
              (apps/adjoint.py:208-217, checkpoints every writeInterval).
  revolve     binomial checkpointing (Griewank & Walther, "Algorithm 799:
              revolve", ACM TOMS 26(1), 2000) with `snapshots` stored states
              beyond the initial one: the minimal number of forward steps
              for that memory.

The schedules are written against a small stepper interface so that the same
code drives the device kernels and the CPU tests:

  stepper.clone(u)              -> an independent copy of state u
  stepper.forward(u, n, keep)   advance u in place over step n; with keep,
                                return a token (the step's tape) that
                                reverse accepts, else None
  stepper.reverse(u_n, n, tok)  adjoint of step n given its start state u_n
                                (the stepper holds the adjoint state); tok
                                None means the step is recomputed from u_n
"""
from math import comb


def beta(c, t):
    """Number of steps that c checkpoints (the initial state included) and t
    forward repetitions can reverse: C(c+t, c)."""
    if c < 0 or t < 0:
        return 0
    return comb(c + t, c)


def repetitions(n, c):
    """Smallest t with beta(c, t) >= n (c >= 1)."""
    t = 0
    while beta(c, t) < n:
        t += 1
    return t


def min_forward_steps(n, snapshots):
    """Forward steps revolve needs to reverse n steps with `snapshots` states
    besides the initial one (Griewank & Walther, Prop. 1, with c = snapshots+1
    checkpoints): t*n - beta(c+1, t-1), t = repetitions(n, c).  Steps re-run
    inside a reverse step (its own stage recomputation) are not counted."""
    if n <= 1:
        return 0
    c = snapshots + 1
    t = repetitions(n, c)
    return t * n - beta(c + 1, t - 1)


def store_all(stepper, u0, n0, n1):
    """Reverse steps [n0, n1) keeping every start state."""
    states = []
    u = stepper.clone(u0)
    for n in range(n0, n1):
        tok = None
        states.append(stepper.clone(u))
        if n + 1 < n1:
            tok = stepper.forward(u, n, True)
        states[-1] = (states[-1], tok)
    for n in range(n1 - 1, n0 - 1, -1):
        un, tok = states.pop()
        stepper.reverse(un, n, tok)


def uniform(stepper, u0, n0, n1, interval, checkpoints=None):
    """Reverse steps [n0, n1) with a state kept every `interval` steps.
    `checkpoints` (dict n -> state) may already hold the segment starts (the
    reference reads them from the written time directories); missing ones
    are computed by one forward sweep."""
    starts = list(range(n0, n1, interval))
    cps = dict(checkpoints or {})
    missing = [k for k in starts if k not in cps]
    if missing:
        u = stepper.clone(u0)
        cps.setdefault(n0, stepper.clone(u0))
        for n in range(n0, max(missing)):
            stepper.forward(u, n, False)
            if n + 1 in missing:
                cps[n + 1] = stepper.clone(u)
    for k in reversed(starts):
        store_all(stepper, cps.pop(k), k, min(k + interval, n1))


def revolve(stepper, u0, n0, n1, snapshots):
    """Binomial reverse of steps [n0, n1) with `snapshots` states stored
    besides u0 (u0 is not modified)."""
    if snapshots < 0:
        raise ValueError('snapshots must be >= 0')
    _revolve(stepper, u0, n0, n1, snapshots)


def _revolve(stepper, ua, a, b, s):
    n = b - a
    if n <= 0:
        return
    if n == 1:
        stepper.reverse(ua, a, None)
        return
    if s == 0:
        # no free snapshot: re-run from a for every step, last step first
        for k in range(b - 1, a - 1, -1):
            u = stepper.clone(ua)
            for m in range(a, k):
                stepper.forward(u, m, False)
            stepper.reverse(u, k, None)
        return
    m = _split(n, s)
    u = stepper.clone(ua)
    for k in range(a, a + m):
        stepper.forward(u, k, False)
    _revolve(stepper, u, a + m, b, s - 1)
    del u
    _revolve(stepper, ua, a, a + m, s)


def _split(n, s):
    """First forward advance m of a binomial reverse of n steps with s free
    snapshots: the right part [m, n) is reversed with s-1 snapshots and at
    most t repetitions, the left part [0, m) with s snapshots and at most t-1,
    and m is taken where the total forward count reaches the minimum."""
    c = s + 1
    t = repetitions(n, c)
    best = min_forward_steps(n, s)
    lo = max(1, n - beta(c - 1, t))
    hi = min(n - 1, beta(c, t - 1))
    for m in range(lo, hi + 1):
        if m + min_forward_steps(n - m, s - 1) + min_forward_steps(m, s) == best:
            return m
    raise AssertionError('no optimal split for n=%d, snapshots=%d' % (n, s))
