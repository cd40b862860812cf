"""CPU model of the bulge-chasing schedule of csrc/sbr.hip (k_sbr_chase).

The kernel reduces a band of half-width B to tridiagonal form (the reduction
inside PDSYEV('N','L') at UTILS.f90:835).  Sweep s annihilates column s below
the subdiagonal; its task k works on rows p = s + 1 + B k .. p + B - 1.  The
kernel runs the sweeps in groups of m consecutive sweeps, one workgroup per
group and one compute wavefront per sweep, in lock step: at step t compute
wave j runs task t - 2 j of sweep s0 + j.  The waves of a group share a ring
buffer of band columns in LDS (the window, WC = 2 B (m + 1) columns).  A separate
loading wave brings the columns of the group's first sweep into the window,
one task ahead, once the previous group has published them; a storing wave,
in the step after each task of the group's last sweep, stores what that task
made final and publishes its progress; the last sweep's wave sends the heads
of its reflectors through mailboxes.

This model runs the same schedule with numpy on the CPU, with
* a race detector on the window: within one phase (between two barriers) no
  element may be written by one actor and read or written by another;
* slot ownership: an actor may touch a window slot only while it holds the
  column the actor expects;
* memory stores that become visible at random times (in issue order per
  wave), but before the progress value that covers them is published, and
  workgroups interleaved at random;
* a check that no value the loading wave uses was in flight from, or stored
  again later by, the previous group (the load could see either value);
and checks the tridiagonal bit for bit against the plain sweep-by-sweep
chase with the same task arithmetic, and its eigenvalues against LAPACK.
"""
import numpy as np
import pytest

def reflector(xn2, alp):
    if xn2 == 0.0:
        return 0.0, alp, 0.0
    beta = -np.copysign(np.sqrt(alp * alp + xn2), alp)
    return (beta - alp) / beta, beta, 1.0 / (alp - beta)


def task_reflector(L, l):
    """Phase A: the reflector of the first column of L (rows < l)."""
    x = L[:, 0]
    xn2 = float(np.sum(x[1:l] ** 2))
    tau, beta, sc = reflector(xn2, float(x[0]))
    v = np.zeros(len(x))
    v[0] = 1.0
    v[1:l] = sc * x[1:l]
    return tau, beta, v


def task_update(L, D, Bn, tau, beta, v):
    """Phase B: H L, H D H and Bn H for H = I - tau v v^T."""
    L = L - tau * np.outer(v, v @ L)
    L[:, 0] = 0.0
    L[0, 0] = beta
    y = tau * (D @ v)
    K = 0.5 * tau * (v @ y)
    w = y - K * v
    D = D - np.outer(v, w) - np.outer(w, v)
    Bn = Bn - tau * np.outer(Bn @ v, v)
    return L, D, Bn


def sizes(n, B, s, k):
    p = s + 1 + B * k
    return p, min(B, n - p), max(0, min(B, n - p - B))


def ntasks(n, B, s):
    return (n - 3 - s) // B + 1 if s <= n - 3 else 0


def gather(get, n, B, p, l, l2):
    D = np.zeros((B, B))
    Bn = np.zeros((B, B))
    for c in range(l):
        for r in range(l):
            D[r, c] = get(p + c, r - c) if r >= c else get(p + r, c - r)
        for r in range(l2):
            Bn[r, c] = get(p + c, B + r - c)
    return D, Bn


def band_of(A, B):
    n = A.shape[0]
    AB = np.zeros((n, 2 * B))
    for c in range(n):
        for d in range(B + 1):
            if c + d < n:
                AB[c, d] = A[c + d, c]
    return AB


def chase_sequential(AB, n, B):
    """Every task of sweep 0, then of sweep 1, ... on one band array."""
    AB = AB.copy()
    for s in range(n - 2):
        L = None
        for k in range(ntasks(n, B, s)):
            p, l, l2 = sizes(n, B, s, k)
            if k == 0:
                L = np.zeros((B, B))
                L[:l, 0] = AB[s, 1:1 + l]
            tau, beta, v = task_reflector(L, l)
            D, Bn = gather(lambda c, d: AB[c, d], n, B, p, l, l2)
            L, D, Bn = task_update(L, D, Bn, tau, beta, v)
            if k == 0:
                AB[s, 1:1 + l] = L[:l, 0]
            else:
                for c in range(B):
                    for r in range(l):
                        AB[p - B + c, B + r - c] = L[r, c]
            for c in range(l):
                for r in range(c, l):
                    AB[p + c, r - c] = D[r, c]
            if k + 1 == ntasks(n, B, s):
                for c in range(l):
                    for r in range(l2):
                        AB[p + c, B + r - c] = Bn[r, c]
            L = Bn
    return AB


class Race(AssertionError):
    pass


LOADER = -1  # actor id of the loading wave
STORER = -2  # actor id of the storing wave


class Window:
    """The LDS ring buffer of one workgroup with per-phase access sets."""

    def __init__(self, WC, LDB):
        self.WC = WC
        self.data = np.zeros((WC, LDB))
        self.owner = [-1] * WC
        self.reads = {}
        self.writes = {}

    def claim(self, col, retired_below):
        slot = col % self.WC
        old = self.owner[slot]
        if old >= 0 and old >= retired_below:
            raise Race(f"slot of column {old} reused for {col} while still live")
        self.owner[slot] = col
        self.data[slot, :] = np.nan  # uninitialised memory

    def _slot(self, col):
        slot = col % self.WC
        if self.owner[slot] != col:
            raise Race(f"column {col} not resident (slot holds {self.owner[slot]})")
        return slot

    def get(self, wave, col, d):
        key = (col, d)
        w = self.writes.get(key)
        if w is not None and w != wave:
            raise Race(f"wave {wave} reads {key} written by wave {w} in the same phase")
        self.reads.setdefault(key, set()).add(wave)
        v = self.data[self._slot(col), d]
        if np.isnan(v):
            raise Race(f"wave {wave} reads uninitialised {key}")
        return v

    def put(self, wave, col, d, v):
        key = (col, d)
        w = self.writes.get(key)
        if w is not None and w != wave:
            raise Race(f"{key} written by waves {w} and {wave} in the same phase")
        rs = self.reads.get(key, set()) - {wave}
        if rs:
            raise Race(f"wave {wave} writes {key} read by waves {rs} in the same phase")
        self.writes[key] = wave
        self.data[self._slot(col), d] = v

    def barrier(self):
        self.reads = {}
        self.writes = {}


class Memory:
    """Band, progress counters and mailboxes in device memory.  A store stays
    pending (invisible to loads) until it is committed; commits happen at
    random, in the order each wave issued its stores.  The loading wave
    publishes a task's progress once its stores are complete."""

    def __init__(self, AB, n, rng, p_commit):
        self.AB = AB.copy()
        self.prog = np.zeros(n, dtype=int)
        self.mail = {}
        self.pending = []  # [owner, seq, kind, key, value]
        self.seq = 0
        self.history = {}  # band element -> [(owner, seq)] of its stores
        self.used = []  # (owner whose later stores would race, element, seq of the load)
        self.rng = rng
        self.p_commit = p_commit

    def store(self, owner, kind, key, v):
        self.pending.append((owner, self.seq, kind, key, v))
        if kind == "ab":
            self.history.setdefault(key, []).append((owner, self.seq))
        self.seq += 1

    def check_loads(self):
        """A value wave 0 loaded and used must not be stored again by the
        group it came from: the load could return either value."""
        for owner, key, seq in self.used:
            for o, sq in self.history.get(key, []):
                if o == owner and sq >= seq:
                    raise Race(f"element {key} loaded at {seq} and stored by group {o} at {sq}")

    def commit_some(self):
        """Completes a random number of each wave's oldest pending stores."""
        owners = sorted({item[0] for item in self.pending})
        for o in owners:
            mine = [item for item in self.pending if item[0] == o]
            cut = int(self.rng.binomial(len(mine), self.p_commit))
            if cut:
                self.complete(o, mine[cut - 1][1] + 1)

    def complete(self, owner, before_seq=None):
        """Commits every pending store of owner issued before before_seq."""
        keep = []
        for item in self.pending:
            if item[0] == owner and (before_seq is None or item[1] < before_seq):
                self._apply(item)
            else:
                keep.append(item)
        self.pending = keep

    def commit_all(self):
        for item in self.pending:
            self._apply(item)
        self.pending = []

    def _apply(self, item):
        _, _, kind, key, v = item
        if kind == "ab":
            self.AB[key] = v
        else:
            self.mail[key] = v


class Group:
    """One workgroup: sweeps s0 .. s0 + m - 1 in lock step."""

    def __init__(self, g, n, B, m, WC, mem, mask=True, dmask=True):
        self.n, self.B, self.m = n, B, m
        self.s0 = g * m
        self.sweeps = [s for s in range(self.s0, self.s0 + m) if s <= n - 3]
        self.last = len(self.sweeps) - 1  # the storing wave
        self.nt = [ntasks(n, B, s) for s in self.sweeps]
        self.steps = max(self.nt[j] + 2 * j for j in range(len(self.sweeps)))
        self.win = Window(WC, 2 * B)
        self.mem = mem
        self.t = -1  # the loading wave's prologue
        self.phase = 0
        self.L = [None] * m
        self.refl = [None] * m
        self.colv = None
        self.colv_mail = None
        self.colv_seq = 0
        self.in_flight = set()
        self.mask = mask
        self.dmask = dmask
        self.ntp = ntasks(n, B, self.s0 - 1) if self.s0 > 0 else 0
        self.done = False

    def retired_below(self):
        """Columns below this are used by no wave of the group from this step
        on (a wave not yet started will still read its own sweep column, and
        the storing wave stores the last sweep's task of the previous step)."""
        low = self.n
        for j, s in enumerate(self.sweeps):
            k = max(self.t, 0) - 2 * j
            if k >= self.nt[j]:
                continue
            if k <= 0:
                low = min(low, s)
            else:
                low = min(low, sizes(self.n, self.B, s, k)[0] - self.B)
        ks = self.t - 1 - 2 * self.last
        if 0 <= ks < self.nt[self.last]:
            low = min(low, self.s0 if ks == 0 else sizes(self.n, self.B, self.sweeps[-1], ks)[0] - self.B)
        return low

    def ready(self):
        """The loading wave's dependencies on the previous group: before the
        first step, the columns of task 0 and its mailbox head; in step t, the
        columns of task t + 1, which it loads and writes into the window then,
        and that task's mailbox head."""
        if self.s0 == 0:
            return True
        t, nt0 = self.t, self.nt[0]
        prog = self.mem.prog[self.s0 - 1]
        if t == -1:
            return prog >= min(2, self.ntp) and (2 >= self.ntp or (self.s0 - 1, 2) in self.mem.mail)
        if t + 1 >= nt0:
            return True
        k = t + 1
        return prog >= min(k + 2, self.ntp) and \
            not (k + 2 < self.ntp and (self.s0 - 1, k + 2) not in self.mem.mail)

    def load_task(self, k):
        """Wave 0's loads for task k: its columns (all 2B diagonals, one
        column per load; at the last task every column up to n - 1) and, if
        it has already arrived, the previous sweep's head for it."""
        n, B, s = self.n, self.B, self.sweeps[0]
        p, l, _ = sizes(n, B, s, k)
        lc = (n - p) if k + 1 == self.nt[0] else l
        # B columns at least; those past n - 1 load as zeros
        self.colv = [self.mem.AB[p + q, :].copy() if p + q < n else np.zeros(2 * B)
                     for q in range(max(lc, B))]
        self.mark_load()
        self.colv_mail = self.mem.mail.get((s - 1, k + 2)) if (s > 0 and k + 2 < self.ntp) else None

    def mark_load(self):
        """At a load: the previous group's stores not yet complete, whose
        elements the load may see with either value."""
        self.colv_seq = self.mem.seq
        prev = self.s0 - self.m
        self.in_flight = {it[3] for it in self.mem.pending if it[0] == prev and it[2] == "ab"}

    def use(self, c, d):
        """Records that wave 0 used the loaded value of (c, d)."""
        if self.s0 > 0:
            if (c, d) in self.in_flight:
                raise Race(f"element {(c, d)} loaded while the previous group's store was in flight")
            self.mem.used.append((self.s0 - self.m, (c, d), self.colv_seq))

    def run_phase(self, rng):
        if self.t == -1:
            self.loader_prologue()
            self.win.barrier()
            self.t = 0
            return
        # one barrier per step: a compute wave forms the reflector of its next
        # task right after its update (its L is the Bn just written), and the
        # reflector of its first task at the start of that task's step
        actors = [(j, self.t - 2 * j) for j in range(len(self.sweeps))]
        actors = [(j, k) for j, k in actors if 0 <= k < self.nt[j]]
        if self.t + 1 < self.nt[0]:
            actors.append((LOADER, self.t + 1))
        ks = self.t - 1 - 2 * self.last  # the last sweep's task of the previous step
        if 0 <= ks < self.nt[self.last]:
            actors.append((STORER, ks))
        order = list(range(len(actors)))
        rng.shuffle(order)
        for i in order:
            j, k = actors[i]
            if j == LOADER:
                self.load_task(k)
                self.put_task(k)
            elif j == STORER:
                self.store(k)
                self.mem.complete(self.s0)
                self.mem.prog[self.sweeps[-1]] = k + 1
            else:
                if k == 0:
                    self.phase_a(j, 0)
                self.phase_b(j, k)
                if k + 1 < self.nt[j]:
                    self.phase_a(j, k + 1)
        self.win.barrier()
        self.t += 1
        if self.t > self.steps:  # the storing wave's last stores run at t = steps
            self.done = True

    def loader_prologue(self):
        """The loading wave before the first step: sweep column s0 and the
        columns of task 0."""
        n, B, s, W, j = self.n, self.B, self.sweeps[0], self.win, LOADER
        col0 = self.mem.AB[s, :].copy()
        self.mark_load()
        W.claim(s, self.retired_below())
        for d in range(2 * B):
            keep = d <= B or not self.mask
            if keep:
                self.use(s, d)
            W.put(j, s, d, col0[d] if keep else 0.0)
        self.load_task(0)
        self.put_task(0)

    def put_task(self, k):
        """Task k's loaded columns into the window (in the step before the
        task's), restricted to the previous sweep's coverage; at the last task
        also the columns past n - 1, whose slots the task owns."""
        n, B, s, W, j = self.n, self.B, self.sweeps[0], self.win, LOADER
        p, l, l2 = sizes(n, B, s, k)
        need_mail = s > 0 and k + 2 < self.ntp
        mailv = self.colv_mail
        if need_mail and mailv is None:
            mailv = self.mem.mail[(s - 1, k + 2)]
        for q, col in enumerate(self.colv):
            W.claim(p + q, self.retired_below())
            vals = col.copy()
            used = np.ones(2 * B, dtype=bool)
            if s > 0 and self.mask:
                # Memory holds the previous sweep's values only inside its
                # coverage; beyond it the band is zero after sweep s - 1,
                # but memory may hold what an earlier group stored there
                # before a later sweep of that group annihilated it.
                if q == B - 1:
                    vals[B + 1:] = 0.0
                    used[B + 1:] = False
                    if need_mail:
                        vals[B] = mailv  # the previous sweep's next head
                        used[B] = False
                else:
                    vals[max(1, 2 * B - 1 - q):] = 0.0
                    used[max(1, 2 * B - 1 - q):] = False
            for d in range(2 * B):
                if used[d] and p + q < n:
                    self.use(p + q, d)
                W.put(j, p + q, d, vals[d])

    def phase_a(self, j, k):
        n, B, s = self.n, self.B, self.sweeps[j]
        p, l, l2 = sizes(n, B, s, k)
        W = self.win
        L = np.zeros((B, B))
        if k == 0:
            for r in range(l):
                L[r, 0] = W.get(j, s, 1 + r)
        else:
            # written as Bn by this wave's previous task, just before
            for c in range(B):
                for r in range(l):
                    L[r, c] = W.get(j, p - B + c, B + r - c)
        self.L[j] = L
        tau, beta, v = task_reflector(self.L[j], l)
        self.refl[j] = (tau, beta, v)
        if k >= 1:
            W.put(j, p - B, B, beta)  # the head (for wave j + 1 from task 2 on)
        if k >= 2:
            if j == self.last:
                self.mem.store(-1 - self.s0, "mail", (s, k), beta)

    def phase_b(self, j, k):
        n, B, s = self.n, self.B, self.sweeps[j]
        p, l, l2 = sizes(n, B, s, k)
        W = self.win
        last = k + 1 == self.nt[j]
        tau, beta, v = self.refl[j]
        D, Bn = gather(lambda c, d: W.get(j, c, d), n, B, p, l, l2)
        L, D, Bn = task_update(self.L[j], D, Bn, tau, beta, v)
        if k == 0:
            for r in range(l):
                W.put(j, s, 1 + r, L[r, 0])
        else:
            for c in range(B):
                for r in range(l):
                    if not (r == 0 and c == 0):  # the head, written with the reflector
                        W.put(j, p - B + c, B + r - c, L[r, c])
        for c in range(l):
            for r in range(c, l):
                W.put(j, p + c, r - c, D[r, c])
        for c in range(l):
            for r in range(l2):
                W.put(j, p + c, B + r - c, Bn[r, c])
        self.L[j] = None

    def store(self, k):
        """The storing wave's stores of the last sweep's task k, in the step
        after the task's.  Whole 16-byte pairs of diagonals: the previous
        block's L columns at diagonals B .. 2B - 1 shifted by the column (head
        included: the next group takes the head from the mailbox and stores
        its own value later) and diagonals 0 .. B - 1 - q of the task's
        column q (the lower half of D: the rest is the task's Bn, which the
        next task is updating in this step and which it stores as that task's
        L), a pair straddling the end alone; at the last task every diagonal
        0 .. B - 1 and its Bn."""
        n, B, W, mem = self.n, self.B, self.win, self.mem
        s = self.sweeps[-1]
        p, l, l2 = sizes(n, B, s, k)
        last = k + 1 == self.nt[self.last]
        rd = lambda c, d: W.get(STORER, c, d)
        st = lambda c, d: mem.store(self.s0, "ab", (c, d), rd(c, d))
        if k == 0:
            for c in range(self.s0, s + 1):
                for d in range(2 if c < s else 1 + l):
                    st(c, d)
        else:
            for c in range(B):
                for r in range(B):
                    st(p - B + c, B + r - c)
        for c in range(l):
            top = B if last or not self.dmask else B - c
            for d in range(top):
                st(p + c, d)
        if last:
            for c in range(l):
                for r in range(l2):
                    st(p + c, B + r - c)
            if p + l == n - 1:
                st(n - 1, 0)  # the last diagonal element, outside this task's block


def chase_grouped(AB, n, B, m, WC, seed, max_active=4, p_commit=0.3, mask=True, dmask=True):
    rng = np.random.default_rng(seed)
    mem = Memory(AB, n, rng, p_commit)
    ngroups = (n - 2 + m - 1) // m if n >= 3 else 0
    groups = []
    nxt = 0
    while True:
        while nxt < ngroups and sum(not g.done for g in groups) < max_active:
            g = Group(nxt, n, B, m, WC, mem, mask, dmask)
            groups.append(g)
            nxt += 1
        active = [g for g in groups if not g.done]
        if not active:
            break
        runnable = [g for g in active if g.ready()]
        if not runnable:
            if not mem.pending:
                raise AssertionError("deadlock")
            mem.commit_some()
            continue
        g = runnable[rng.integers(len(runnable))]
        g.run_phase(rng)
        if rng.random() < 0.5:
            mem.commit_some()
    mem.commit_all()
    mem.check_loads()
    return mem.AB


def window_cols(B, m):
    """The kernel's window: the columns in use span 2 B m - m + 2 + B (the
    storing wave stores the last sweep's previous task), and the loading
    wave writes the next task's B columns a step ahead."""
    return max(2 * B * (m + 1), 2 * B * m - m + 2 + 2 * B)


def random_band(n, B, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    A = A + A.T
    i, j = np.indices((n, n))
    A[np.abs(i - j) > B] = 0.0
    return A


CASES = [(B, m, n) for B, m, n in [
    (4, 4, 3), (4, 4, 5), (4, 4, 9), (4, 4, 23), (4, 4, 40), (4, 4, 61),
    (4, 2, 37), (4, 3, 44), (4, 1, 30), (3, 4, 50), (5, 3, 47), (8, 4, 90),
    (32, 4, 100), (32, 3, 130), (32, 3, 200),
]]


@pytest.mark.parametrize("B,m,n", CASES)
def test_grouped_chase_matches_sequential(B, m, n):
    A = random_band(n, B, 7 + n)
    AB = band_of(A, B)
    ref = chase_sequential(AB, n, B)
    for seed in range(3 if B < 32 else 1):
        got = chase_grouped(AB, n, B, m, window_cols(B, m), seed)
        assert np.array_equal(got[:, 0], ref[:, 0])
        assert np.array_equal(got[:n - 1, 1], ref[:n - 1, 1])
    d, e = ref[:, 0], ref[:n - 1, 1]
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w = np.linalg.eigvalsh(T)
    wr = np.linalg.eigvalsh(A)
    assert np.abs(w - wr).max() <= 1e-12 * np.abs(wr).max() * n


def test_first_wave_must_zero_beyond_coverage():
    # Without zeroing the diagonals beyond the previous sweep's coverage the
    # first wave uses elements the previous group is still storing, and the
    # group's later sweeps read values an earlier group stored before a sweep
    # inside a later group's window annihilated them.
    B, m, n = 4, 2, 37
    A = random_band(n, B, 7 + n)
    AB = band_of(A, B)
    ref = chase_sequential(AB, n, B)
    wrong = 0
    for seed in range(4):
        try:
            got = chase_grouped(AB, n, B, m, window_cols(B, m), seed, mask=False)
            wrong += not np.array_equal(got[:, 0], ref[:, 0])
        except Race:
            wrong += 1
    assert wrong == 4


def test_window_too_small_is_caught():
    B, m, n = 4, 4, 61
    AB = band_of(random_band(n, B, 3), B)
    with pytest.raises(Race):
        chase_grouped(AB, n, B, m, 2 * B * m, 0)


def test_store_must_stop_before_the_next_tasks_rows():
    # Storing diagonals B - q .. B - 1 of the task's column q with its D half
    # reads the next task's L while that task is updating it.
    B, m, n = 4, 3, 44
    AB = band_of(random_band(n, B, 7 + n), B)
    with pytest.raises(Race):
        chase_grouped(AB, n, B, m, window_cols(B, m), 0, dmask=False)
