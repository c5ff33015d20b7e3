"""CPU model of the partner lists' partial builds (the rule the frame kernels
follow), checked against brute force: in every frame, each live element's
list holds every element it touches.

The model keeps what the kernels keep per element (list, build position,
activity) and per coarse cell (the stamp: 1 + the build count when a live
element stood there), and applies the kernels' rules:

* a frame builds when some element of the previous frame ended farther than
  the threshold (skin / 2 - 5 mm) from its build position, and a call's
  first frame builds when any element lies farther than that (or on the
  first call); a call's first-frame build rebuilds every list;
* a later build rebuilds the lists of live elements and of frozen elements
  whose reach box (2 r_max + 2 delta + skin) holds a stamp >= the build count
  seen by the frame before;
* a live element stamps its end cell on a build frame (and its build cell),
  on its activation frame, and when its end cell differs from its build
  position's;
* the call-start cull stamps the cell of every active element and of every
  element away from its build position.

Elements move by random steps within a frame (live ones only), activate at
random, and the caller moves and flips elements between calls.  A scripted
case then walks a frozen element toward a frozen neighbour in caller moves
under the threshold while a live element beside it (out of the neighbour's
reach box) keeps the builds going: with the cull's stamps the neighbour's
list takes it in; without them (a mistake planted in the model) the
neighbour, activated, misses the contact, which shows the check can fail.
"""
import numpy as np
import pytest

DELTA = 0.05
SKIN = 1.0
THR = SKIN / 2 - 0.005
CELL2 = 1.5
BOX = 48.0


class Model:
    def __init__(self, rng, n=150, cull_stamps=True, pos=None, r=None):
        self.rng = rng
        self.n = n if pos is None else len(pos)
        n = self.n
        self.r = rng.uniform(0.4, 1.2, n) if r is None else np.asarray(r, float)
        self.pos = rng.uniform(0.0, BOX, (n, 2)) if pos is None else np.asarray(pos, float)
        self.act = np.zeros(n, bool)
        self.xb = np.full((n, 2), np.nan)
        self.lists = [np.zeros(0, int)] * n
        self.nc = int(np.ceil(BOX / CELL2)) + 2
        self.stamp = np.zeros((self.nc, self.nc), np.int64)
        self.count = 0
        self.reach = (2 * self.r.max() + 2 * DELTA + SKIN) * 1.001 + 1e-3
        self.cull_stamps = cull_stamps
        self.first = True

    def cell(self, p):
        c = np.floor((np.asarray(p) + CELL2) / CELL2).astype(int)
        return np.clip(c, 0, self.nc - 1)

    def build(self, idx):
        d = np.linalg.norm(self.pos[idx][:, None, :] - self.pos[None, :, :], axis=2)
        lim = self.r[idx][:, None] + self.r[None, :] + 2 * DELTA + SKIN
        for a, e in enumerate(idx):
            m = d[a] <= lim[a]
            m[e] = False
            self.lists[e] = np.flatnonzero(m)
            self.xb[e] = self.pos[e]

    def stamped_near(self, e):
        lo = self.cell(self.pos[e] - self.reach)
        hi = self.cell(self.pos[e] + self.reach)
        return bool((self.stamp[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1] >= self.count).any())

    def put(self, p, v):
        c = self.cell(p)
        self.stamp[c[0], c[1]] = max(self.stamp[c[0], c[1]], v)

    def check_contacts(self):
        live = np.flatnonzero(self.act)
        if len(live) == 0:
            return 0
        d = np.linalg.norm(self.pos[live][:, None, :] - self.pos[None, :, :], axis=2)
        lim = self.r[live][:, None] + self.r[None, :] + 2 * DELTA
        missed = 0
        for a, e in enumerate(live):
            touch = np.flatnonzero(d[a] <= lim[a])
            touch = touch[touch != e]
            missed += len(np.setdiff1d(touch, self.lists[e]))
        return missed

    def call(self, frames, step=0.08, p_act=0.01, motion=None):
        """One call of `frames` frames; returns the contacts missed.
        motion(k, live) gives the live elements' steps in place of random
        ones (and then nothing activates)."""
        far = self.first or bool((np.linalg.norm(self.pos - self.xb, axis=1) > THR).any())
        self.first = False
        if self.cull_stamps:
            for e in np.flatnonzero(self.act | (self.pos != self.xb).any(axis=1)):
                self.put(self.pos[e], self.count + 1)
        missed = 0
        rebuild_next = False
        need = np.ones(self.n, bool)
        for k in range(frames):
            rebuild = far if k == 0 else rebuild_next
            act_raw = self.act.copy()
            # activation at the frame's start (the element is live this frame)
            if motion is None:
                self.act |= self.rng.random(self.n) < p_act
            if rebuild:
                self.build(np.arange(self.n) if k == 0 else np.flatnonzero(self.act | need))
            missed += self.check_contacts()
            # motion of the live elements, the build test and the stamps
            start = self.pos.copy()
            live = np.flatnonzero(self.act)
            self.pos[live] += (self.rng.uniform(-step, step, (len(live), 2)) if motion is None
                               else motion(k, live))
            st = self.count + 1      # the count before this frame's build is counted
            for e in live:
                cn, cb = self.cell(self.pos[e]), self.cell(self.xb[e])
                if rebuild or not act_raw[e] or (cn != cb).any():
                    self.put(self.pos[e], st)
                    if rebuild:
                        self.put(start[e], st)
            if rebuild:
                self.count += 1
            disp = np.linalg.norm(self.pos[live] - self.xb[live], axis=1)
            rebuild_next = bool((disp > THR).any())
            if rebuild_next:   # the frame before a build marks the frozen elements to rebuild
                need = self.act.copy()
                for e in np.flatnonzero(~self.act):
                    need[e] = self.stamped_near(e)
        return missed

    def caller_edit(self):
        """Moves (mostly within the threshold) and activity flips."""
        rng = self.rng
        pick = rng.choice(self.n, 12, replace=False)
        ang = rng.uniform(0, 2 * np.pi, len(pick))
        size = np.where(rng.random(len(pick)) < 0.9, 0.9 * THR, 1.5 * THR)
        self.pos[pick] += size[:, None] * np.stack([np.cos(ang), np.sin(ang)], axis=1)
        flip = rng.choice(self.n, 4, replace=False)
        self.act[flip] = ~self.act[flip]


def run(seed, calls=20):
    rng = np.random.default_rng(seed)
    m = Model(rng)
    missed = 0
    for c in range(calls):
        if c:
            m.caller_edit()
        missed += m.call(int(rng.integers(3, 25)))
    return missed, m.count


@pytest.mark.parametrize("seed", range(3))
def test_partial_builds_keep_every_contact(seed):
    missed, builds = run(seed)
    assert builds > 5
    assert missed == 0


def scripted(cull_stamps):
    """e (0) and j (1) frozen just beyond list reach; k (2) live beside j,
    outside e's reach box.  Each call k steps past the threshold and back
    (one build per call, every element within the threshold at the call's
    end) and the caller then moves j 0.9 threshold toward e; after four moves
    they touch, and the caller activates e."""
    r = 0.8
    gap = 2 * r + 2 * DELTA + SKIN + 0.1
    pos = [[10.0, 10.0], [10.0 + gap, 10.0], [10.0 + gap + 2.2, 10.0]]
    m = Model(np.random.default_rng(0), cull_stamps=cull_stamps, pos=pos, r=[r, r, r])
    m.act[2] = True
    missed = 0
    for c in range(7):
        if c:
            m.pos[1, 0] -= 0.9 * THR
        if c == 5:
            m.act[0] = True
        sgn = 1.0 if c % 2 == 0 else -1.0

        def motion(k, live):
            d = np.zeros((len(live), 2))
            d[live == 2, 0] = sgn * (0.6 if k == 0 else -0.3 if k == 1 else 0.0)
            return d
        missed += m.call(3, motion=motion)
    return missed, m.count


def test_scripted_caller_moves():
    missed, builds = scripted(True)
    assert builds >= 7 and missed == 0


def test_scripted_caller_moves_without_cull_stamps_misses():
    # the planted mistake: caller edits leave no stamp
    missed, builds = scripted(False)
    assert builds >= 7 and missed > 0
