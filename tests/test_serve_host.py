"""Host-side invariants of the server's thread plumbing (CPU, with stand-in
device objects): graphs are created and destroyed only under the server's
capture lock, so no graph is destroyed while another thread captures (the
abort 'operation not permitted when stream is capturing' from a graph
destructor), and a finished solve's buffers are returned only after its
build stream has waited for its solve stream."""
import threading

import numpy as np
import pytest

from tectosaur_amd import serve


class _LockProbe:
    """A lock that records whether it was held when asked."""

    def __init__(self):
        self._lk = threading.Lock()

    def __enter__(self):
        self._lk.acquire()
        return self

    def __exit__(self, *a):
        self._lk.release()

    def held(self):
        return self._lk.locked()


class _DelProbe:
    def __init__(self, lock, log):
        self.lock, self.log = lock, log

    def __del__(self):
        self.log.append(self.lock.held())


class _Stream:
    def __init__(self, log, name):
        self.log, self.name = log, name

    def wait_stream(self, other):
        self.log.append(('wait', self.name, other.name))


def test_release_destroys_graphs_under_the_capture_lock_after_stream_join():
    lock = _LockProbe()
    held = []
    order = []
    a = serve._Active(None, _DelProbe(lock, held), _DelProbe(lock, held),
                      _DelProbe(lock, held), 0, stream=_Stream(order, 'solve'),
                      build_stream=_Stream(order, 'build'))
    serve._release(a, lock)
    assert held == [True, True, True]       # every object dropped with the lock held
    assert order == [('wait', 'build', 'solve')]
    assert a.op is None and a.state is None and a.prob is None
    assert not lock.held()


def test_adopted_build_is_destroyed_by_its_release_only():
    # the serving loop adopts finished builds through Server._adopt_built:
    # after it returns, only the active solve holds the build's objects, so
    # nothing is destroyed until the release, and the release destroys them
    # all under the capture lock
    import collections
    import concurrent.futures
    lock = _LockProbe()
    held = []

    class Prob(_DelProbe):
        n_tris, pair_evals = 512, 0

    class State(_DelProbe):
        batch_split = False

    class Rec:
        admitted = 0.0

    srv = serve.Server.__new__(serve.Server)
    srv.streams = []
    srv.group_stream = None
    srv._next_stream = 0
    fut = concurrent.futures.Future()
    fut.set_result((Prob(lock, held), _DelProbe(lock, held), State(lock, held), None, None))
    building = collections.deque([(fut, Rec(), 5)])
    del fut
    active = []
    srv._adopt_built(active, building.popleft(), lambda: 1.0)
    assert held == [] and len(active) == 1
    serve._release(active[0], lock)
    assert held == [True, True, True]
    assert not lock.held()
    # the pattern it replaces: the build unpacked in the caller's frame
    # outlives the release, and dropping those names destroys it unlocked
    held.clear()
    fut = concurrent.futures.Future()
    fut.set_result((Prob(lock, held), _DelProbe(lock, held), State(lock, held), None, None))
    prob, op, state, ev, bst = fut.result()
    del fut
    active = []
    srv._adopt(active, Rec(), prob, op, state, 5, lambda: 1.0, ev, bst)
    serve._release(active[0], lock)
    assert held == []
    del prob, op, state
    assert held == [False, False, False]


def test_release_empties_the_allocator_cache_beyond_the_slack(monkeypatch):
    # destroyed graphs leave their memory pools reserved: a release empties
    # the cache, under the capture lock, once the unallocated reserved
    # memory exceeds the slack, and not before
    lock = _LockProbe()
    calls = []
    mem = {'reserved': 0, 'allocated': 0}
    monkeypatch.setattr(serve.torch.cuda, 'memory_reserved', lambda *a, **k: mem['reserved'])
    monkeypatch.setattr(serve.torch.cuda, 'memory_allocated', lambda *a, **k: mem['allocated'])
    monkeypatch.setattr(serve.torch.cuda, 'empty_cache', lambda: calls.append(lock.held()))
    for reserved, allocated, slack, want in ((100, 10, 50, [True]), (100, 60, 50, []),
                                             (10 ** 12, 0, None, [])):
        mem.update(reserved=reserved, allocated=allocated)
        calls.clear()
        a = serve._Active(None, object(), object(), object(), 0)
        serve._release(a, lock, slack)
        assert calls == want, (reserved, allocated, slack)
        assert a.op is None and not lock.held()


def test_build_captures_under_the_capture_lock(monkeypatch):
    lock = _LockProbe()
    seen = {}

    class Farfield:
        tree = object()

    class H:
        farfield = Farfield()

    class Prob:
        def __init__(self, *a, **kw):
            self.H, self.cm, self.rhs = H(), None, None
            self.n_tris = 500   # below the size up to which N-bodies are batched

    class CmT:
        def dot_dev(self, x):
            return x

    class Op:
        def __init__(self, H, cm):
            self.cmT = CmT()

        def enable_graphs(self, precs, stream=None):
            seen['held'] = lock.held()
            seen['precs'] = precs

        def enable_split_graphs(self, precs, stream=None):
            seen['split_held'] = lock.held()
            seen['split_precs'] = precs

    class State:
        batch_split = False

    monkeypatch.setattr(serve, 'TractionSlipProblem', Prob)
    monkeypatch.setattr(serve, 'ConstrainedOperator', Op)
    monkeypatch.setattr(serve, 'make_state', lambda *a, **kw: State())
    _, _, st = serve._build({}, 'tree', 1e-8, 100, 1000, graphs=True, capture_stream=None,
                            capture_lock=lock)
    assert seen['held'] is True and 'split_held' not in seen and not st.batch_split
    assert not lock.held()
    # a direct-farfield solve of a batching server: the inner precision's
    # apply is captured as split graphs, also under the lock
    Farfield.tree = None
    _, _, st = serve._build({}, 'tree', 1e-8, 100, 1000, graphs=True, capture_stream=None,
                            capture_lock=lock, batch_nbody=True)
    assert seen['split_held'] is True and seen['split_precs'] == ('f32',)
    assert 'f32' not in seen['precs'] and st.batch_split
    assert not lock.held()
    # a larger direct-farfield solve keeps whole graphs (its own N-body launch)
    seen.clear()
    big = Prob
    monkeypatch.setattr(serve, 'TractionSlipProblem',
                        lambda *a, **kw: _with(big(), n_tris=serve.NBODY_BATCH_MAX_TRIS))
    _, _, st = serve._build({}, 'tree', 1e-8, 100, 1000, graphs=True, capture_stream=None,
                            capture_lock=lock, batch_nbody=True)
    assert not st.batch_split and 'split_precs' not in seen and 'f32' in seen['precs']


def _with(obj, **kw):
    for k, v in kw.items():
        setattr(obj, k, v)
    return obj


def test_estimate_bytes_grows_with_size_and_tree():
    a = serve.estimate_bytes(1000, 1500)
    b = serve.estimate_bytes(4000, 1500)
    assert 0 < a < b
    assert serve.estimate_bytes(4000, 1500, tree=True) > b
