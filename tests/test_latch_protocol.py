"""Host replay of the job loop of the tree kernels (subtree_insert_kernel and
subtree_summary_kernel in nns_amd/csrc/k_tree.hip, which stride over their
jobs; they took them from a shared counter before): one workgroup of several waves takes jobs
through an LDS word that only thread 0 (in wave 0) writes, while another
workgroup of the same launch may raise the error latch at any moment.

Every interleaving of the waves' steps with the moment the latch is raised is
replayed.  A barrier completes once every wave that has not left has arrived
(the hardware's rule: a wave that ended no longer counts).  LDS starts with a
stale value left by an earlier workgroup on the CU.

  entry check (the earlier code): each wave reads the latch on entry and leaves
      on its own; if wave 0 leaves and the others do not, they pass the barrier
      and use the stale LDS word as a job number.  The replay finds that case.
  uniform decision (the kernels now): thread 0 reads the latch when it takes a
      job and publishes "stop" in the same LDS word; no wave ever uses a value
      that was not written in the same round, and all waves leave together.
"""
import itertools

STALE = -12345  # LDS contents left by a previous workgroup
NSUB = 2        # jobs in the queue


class Sim:
    def __init__(self, nwaves, uniform):
        self.nwaves = nwaves
        self.uniform = uniform
        self.latch = False
        self.lds = STALE
        self.written = False  # LDS word written in the current round
        self.next_job = 0
        self.used = []        # (wave, job) pairs the waves went on to process
        self.bad = []         # uses of a job number nobody wrote this round

    def program(self, w):
        """Yields once per step; yields 'barrier' at a workgroup barrier."""
        if not self.uniform:
            yield "step"
            if self.latch:
                return  # the earlier per-wave entry check
        while True:
            if w == 0:
                yield "step"
                if self.uniform and self.latch:
                    self.lds = NSUB
                else:
                    self.lds = self.next_job
                    self.next_job += 1
                self.written = True
            yield "barrier"
            job = self.lds
            if not self.written:
                # on the GPU: subcells[job] / sub[job] read at a stale index, or
                # the same job taken again and again; the replay stops this wave
                self.bad.append((w, job))
                return
            yield "barrier"
            self.written = False if w == 0 else self.written
            if job >= NSUB:
                return
            self.used.append((w, job))
            yield "step"


def replay(nwaves, uniform, order, raise_at):
    """Runs the waves by `order` (a sequence of wave ids, the next wave to step
    when it can), raising the latch before global step `raise_at`."""
    sim = Sim(nwaves, uniform)
    progs = {w: sim.program(w) for w in range(nwaves)}
    state = {w: "run" for w in range(nwaves)}  # run | wait | done
    step = 0
    oi = 0

    def advance(w):
        try:
            r = next(progs[w])
        except StopIteration:
            state[w] = "done"
            return
        state[w] = "wait" if r == "barrier" else "run"

    while any(s != "done" for s in state.values()):
        live = [w for w in range(nwaves) if state[w] != "done"]
        if all(state[w] == "wait" for w in live):  # barrier complete
            for w in live:
                state[w] = "run"
        runnable = [w for w in live if state[w] == "run"]
        if step == raise_at:
            sim.latch = True
        w = order[oi % len(order)] if order[oi % len(order)] in runnable else runnable[0]
        oi += 1
        advance(w)
        step += 1
        if step > 200:
            raise AssertionError("no progress")
    return sim


def all_cases(nwaves):
    for order in itertools.product(range(nwaves), repeat=4):
        for raise_at in range(0, 12):
            yield order, raise_at


def test_entry_check_lets_waves_use_an_unwritten_job():
    found = []
    for order, raise_at in all_cases(3):
        sim = replay(3, False, order, raise_at)
        if sim.bad:
            found.append((order, raise_at, sim.bad))
    assert found, "the replay should find the case where wave 0 left alone"
    # the case: the latch rises after waves 1 and 2 passed their entry check and
    # before wave 0 made its own; waves 1, 2 then read the stale word
    assert any(any(job == STALE for _, job in bad) for _, _, bad in found)


def test_uniform_decision_never_uses_an_unwritten_job():
    for nwaves in (2, 3):
        for order, raise_at in all_cases(nwaves):
            sim = replay(nwaves, True, order, raise_at)
            assert not sim.bad, (order, raise_at, sim.bad)
            # every wave processed exactly the jobs wave 0 took, in rounds
            jobs = sorted({j for _, j in sim.used})
            assert jobs == list(range(len(jobs)))
            for j in jobs:
                assert sorted(w for w, k in sim.used if k == j) == list(range(nwaves))
