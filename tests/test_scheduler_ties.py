"""Events of equal time.

The reference's heap scheduler ranks equal times by heap.c's structural order (whatever the
sift steps leave at the root, heap.c:156-276); its list scheduler by first push
(list_scheduler.py:129-156).  The device chain keeps, per tagger, the earliest pushed event
and ranks equal times by push order under either scheduler.  The oracle has both rules
(ORA_TIES_SCHEDULER, ORA_TIES_FIRST_PUSH).

- CPU: the oracle's heap restatement equals the reference's own heap.c (compiled into
  oracle/_ref) on tie-heavy sequences of pushes, lazy trashes and root requests; its
  first-push rule equals a brute-force minimum over (time, push order) of the valid entries.
- CPU: on a run with exact ties (sampling events and chain ends both every 0.5), the two rules
  commit different sequences under the heap scheduler, and the same under the list scheduler.
- GPU: on that run the device equals the oracle's first-push rule under both schedulers.
"""
import ctypes

import numpy as np
import pytest

import systems as S


def _tie_ops(rng, n_ops=4000, n_handlers=12):
    op = rng.choice([0, 0, 0, 1, 2], n_ops).astype(np.int32)
    h = rng.integers(0, n_handlers, n_ops).astype(np.int32)
    q = rng.integers(0, 3, n_ops).astype(np.float64)
    r = rng.choice([0.0, 0.25, 0.5], n_ops)
    return op, h, q, r


def _oracle_replay(oracle, op, h, q, r, n_handlers, first_push):
    out = np.zeros(len(op), dtype=np.int32)
    pi = ctypes.POINTER(ctypes.c_int32)
    pd = ctypes.POINTER(ctypes.c_double)
    assert oracle.lib().ora_heap_replay(len(op), op.ctypes.data_as(pi), h.ctypes.data_as(pi), q.ctypes.data_as(pd),
                                        r.ctypes.data_as(pd), n_handlers, int(first_push), out.ctypes.data_as(pi)) == 0
    return out


class _HeapEntry(ctypes.Structure):
    _fields_ = [("time_quotient", ctypes.c_double), ("time_remainder", ctypes.c_double),
                ("event_handler", ctypes.c_void_p), ("counter", ctypes.c_uint)]


def _reference_replay(ref, op, h, q, r, n_handlers):
    """The same operations through the reference's heap.c, with heap_scheduler.py's counters."""
    CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint)
    ref.construct_heap.restype = ctypes.c_void_p
    ref.insert.restype = ctypes.c_size_t
    ref.insert.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_uint]
    ref.root.restype = _HeapEntry
    ref.root.argtypes = [ctypes.c_void_p, ctypes.c_void_p, CB]
    ref.destroy_heap.argtypes = [ctypes.c_void_p]
    counter = [0] * n_handlers

    def stale(_scheduler, handler, cnt):
        return 1 if counter[handler - 1] > cnt else 0
    cb = CB(stale)
    heap = ref.construct_heap()
    out = np.zeros(len(op), dtype=np.int32)
    try:
        for i in range(len(op)):
            out[i] = -2
            if op[i] == 0:
                ref.insert(heap, q[i], r[i], int(h[i]) + 1, counter[h[i]])
            elif op[i] == 1:
                counter[h[i]] += 1
            else:
                e = ref.root(heap, None, cb)
                out[i] = (e.event_handler - 1) if e.event_handler else -1
    finally:
        ref.destroy_heap(heap)
    return out


def test_oracle_heap_equals_reference_heap_c_on_ties(oracle):
    ref = oracle.ref_lib("heap")
    if ref is None:
        pytest.skip("reference C sources not compiled (no reference tree)")
    for seed in range(6):
        op, h, q, r = _tie_ops(np.random.default_rng(seed))
        want = _reference_replay(ref, op, h, q, r, 12)
        got = _oracle_replay(oracle, op, h, q, r, 12, first_push=False)
        assert np.array_equal(got, want), seed
        assert np.sum(want >= 0) > 500


def test_oracle_first_push_ties_are_the_earliest_pushed_valid_entry(oracle):
    for seed in range(6):
        op, h, q, r = _tie_ops(np.random.default_rng(100 + seed))
        got = _oracle_replay(oracle, op, h, q, r, 12, first_push=True)
        counter = [0] * 12
        entries = []   # (q, r, push, handler, counter)
        for i in range(len(op)):
            if op[i] == 0:
                entries.append((q[i], r[i], i, int(h[i]), counter[h[i]]))
            elif op[i] == 1:
                counter[h[i]] += 1
            else:
                valid = [e for e in entries if e[4] >= counter[e[3]]]
                want = min(valid)[3] if valid else -1
                assert got[i] == want, (seed, i)
        structural = _oracle_replay(oracle, op, h, q, r, 12, first_push=False)
        assert not np.array_equal(got[op == 2], structural[op == 2])   # the rules do differ on these inputs


def tie_heavy_system(scheduler):
    """64 hard disks at packing fraction 0.5 with chains of 0.5 and sampling every 0.5: every
    chain end coincides with a sampling event, to the last bit (both are sums of 0.5)."""
    meta, pos = S.hard_disk_system(n=64, phi=0.5, seed=3, chain_time=0.5)
    meta["sampling_interval"] = 0.5
    meta["scheduler"] = scheduler
    return meta, pos


@pytest.mark.parametrize("scheduler", ["heap", "list"])
def test_tie_rules_on_a_run_with_exact_ties(oracle, scheduler):
    meta, pos = tie_heavy_system(scheduler)
    logs = []
    for ties in (oracle.TIES_SCHEDULER, oracle.TIES_FIRST_PUSH):
        c = S.oracle_config(meta, math_mode=oracle.MATH_CR, scheduler=scheduler)
        c.ties = ties
        logs.append(oracle.Engine(c, pos).run(5000))
    a, b = logs
    t = a["tq"] + a["tr"]
    assert np.sum(np.diff(t) == 0) > 100                     # exact ties occur
    same = np.array_equal(a["tagger"], b["tagger"]) and np.array_equal(a["target"], b["target"])
    assert same == (scheduler == "list")


@pytest.mark.gpu
@pytest.mark.parametrize("scheduler", ["heap", "list"])
def test_gpu_chain_ranks_ties_by_push_order(device, oracle, scheduler):
    from jellyfysh_amd.chain import DeviceChain
    meta, pos = tie_heavy_system(scheduler)
    legs = 20000
    print("tie-heavy hard disks, %s scheduler, %d legs" % (scheduler, legs), flush=True)
    got = DeviceChain(S.device_desc(meta, scheduler=scheduler), pos).run(legs)
    c = S.oracle_config(meta, math_mode=oracle.MATH_CR, scheduler=scheduler)
    c.ties = oracle.TIES_FIRST_PUSH
    want = oracle.Engine(c, pos).run(legs)
    for f in ("tagger", "flags", "active", "target", "tq", "tr", "pos", "ncand", "chash"):
        eq = got[f] == want[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.flatnonzero(~eq)
        assert bad.size == 0, (f, int(bad[0]))
