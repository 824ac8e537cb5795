"""Set-level event handlers: the reference's mediator loop on the host, the candidate sets on
the device.

Each tagger of a run becomes one ``CandidateSetEventHandler``: its ``send_event_time`` evaluates
the whole candidate set of the tagger for the active unit on the device (ecmc_system_send_event_time)
and returns ``(Time, [target])``, the form of the reference's handlers that stand for a set of
factors (event_handler/event_handler.py:80-125, event_handler/abstracts/cell_veto_event_handler.py:
162-238).  The mediator then calls ``get_arguments_candidate_set_event_handler(target)`` to build the
arguments of ``send_out_state`` (mediator/mediator.py:142-211, as mediator.py:265-292 does for the
cell veto), and ``send_out_state`` commits the event on the device
(ecmc_system_send_out_state: thinning, lifting, position write-back).

``StepMediator`` is the reference's SingleProcessMediator.run (mediator/single_process_mediator.py:
84-156) over these handlers, with the TagActivator's create and trash lists by tag
(activator/tag_activator.py:166-316) and a list scheduler (scheduler/list_scheduler.py:116-175):
the state, the cell occupancy and the random stream stay on the device, the loop and the
scheduler run in Python.  It commits the same events as the device's own leg loop
(DeviceChain.run) and as the oracle.
"""
import ctypes
import math

import numpy as np

from . import _lib as L
from .chain import DeviceChain


class Time:
    """base/time.py:27-259: a time as quotient and remainder of its division by one, ordered
    lexicographically."""
    __slots__ = ("quotient", "remainder")

    def __init__(self, quotient, remainder):
        self.quotient = float(quotient)
        self.remainder = float(remainder)

    @classmethod
    def from_float(cls, value):
        q, r = divmod(value, 1.0)
        return cls(q, r)

    def __float__(self):
        return self.quotient + self.remainder

    def _key(self):
        return (self.quotient, self.remainder)

    def __lt__(self, other):
        return self._key() < other._key()

    def __le__(self, other):
        return self._key() <= other._key()

    def __eq__(self, other):
        return isinstance(other, Time) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())

    def __repr__(self):
        return "Time(%r, %r)" % (self.quotient, self.remainder)


class CandidateSetEventHandler:
    """The factors of one tagger as one event handler whose candidate set is evaluated on the
    device: send_event_time(*in_state) -> (Time, [target]); send_out_state(target) -> the
    committed event (a record of jellyfysh_amd._lib.EVENT_DTYPE)."""

    number_send_event_time_arguments = 1
    number_send_out_state_arguments = 1

    def __init__(self, system, tagger, tag=None):
        self._system = system
        self.tagger = int(tagger)
        self.tag = tag

    def send_event_time(self, *in_state):
        q, r = ctypes.c_double(), ctypes.c_double()
        target, valid = ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib().ecmc_system_send_event_time(self._system._h, self.tagger, ctypes.byref(q), ctypes.byref(r),
                                                    ctypes.byref(target), ctypes.byref(valid)))
        if not valid.value:
            return Time(math.inf, math.inf), [-1]
        return Time(q.value, r.value), [target.value]

    def send_out_state(self, target):
        """Commit the event whose time this handler sent last.  ``target`` is the candidate that
        send_event_time returned (a unit, or a cell for the cell veto); the device holds the same
        candidate in the handler's slot and resolves it at the commit (a cell to its occupant,
        a thinning to a lift or a rejection), so the committed record's target can differ from it
        (start, sampling and end-of-chain events carry -1)."""
        ev = np.zeros(1, dtype=L.EVENT_DTYPE)
        L.check(L.lib().ecmc_system_send_out_state(self._system._h, self.tagger, ev.ctypes.data))
        return ev[0]


class ListScheduler:
    """scheduler/list_scheduler.py:116-175: the earliest stored time, the first pushed among equal
    times; trash removes the handler's entry."""

    def __init__(self):
        self._times = []
        self._last = None

    def push_event(self, time, event_handler):
        self._times.append((time, event_handler))

    def get_succeeding_event(self):
        if not self._times:
            raise RuntimeError("SchedulerError: the scheduler holds no events")
        best = min(self._times, key=lambda e: e[0])
        if self._last is not None and best[0] < self._last:
            raise RuntimeError("SchedulerError: event time decreased")
        self._last = best[0]
        return best[1]

    def trash_event(self, event_handler):
        for i, (_, h) in enumerate(self._times):
            if h is event_handler:
                del self._times[i]
                return
        raise RuntimeError("SchedulerError: event handler not present in scheduler")


class StepMediator:
    """single_process_mediator.py:84-156 on the host over one CandidateSetEventHandler per tagger.

    ``taggers`` is the tagger list of the run (the dictionaries of jellyfysh_amd.chain.make_desc:
    tag, kind, create, trash); the system is a DeviceChain built from the same description."""

    def __init__(self, system, taggers):
        if not isinstance(system, DeviceChain):
            raise TypeError("a DeviceChain holds the state")
        self.system = system
        self.tags = [t["tag"] for t in taggers]
        self.handlers = [CandidateSetEventHandler(system, i, t["tag"]) for i, t in enumerate(taggers)]
        index = {tag: i for i, tag in enumerate(self.tags)}
        self._create = [[index[x] for x in t.get("create", [])] for t in taggers]
        self._trash = [[index[x] for x in t.get("trash", [])] for t in taggers]
        starts = [i for i, t in enumerate(taggers) if t["kind"] == "start_of_run"]
        if len(starts) != 1:
            raise ValueError("one start-of-run tagger")
        self._start = starts[0]
        self.scheduler = ListScheduler()
        self._out_state_arguments = {}
        self._running = set()      # taggers whose event sits in the scheduler
        self._preceding = None

    def get_arguments_candidate_set_event_handler(self, target):
        """mediator.py's get_arguments_<class> hook: the target travels as the argument."""
        return (target,)

    def run(self, max_events):
        """Commit up to max_events legs (or until the end of run); returns their records."""
        events = []
        done = False
        while len(events) < max_events and not done:
            if self._preceding is None:
                to_run = [self._start]
            else:
                L.check(L.lib().ecmc_system_begin_leg(self.system._h))
                to_run = self._create[self._preceding]
            for t in to_run:
                h = self.handlers[t]
                if t in self._running:
                    raise RuntimeError("TagActivatorError: tagger %s creates while its event is pending" % h.tag)
                event_time, self._out_state_arguments[h] = h.send_event_time()
                self.scheduler.push_event(event_time, h)
                self._running.add(t)
            h = self.scheduler.get_succeeding_event()
            ev = h.send_out_state(*self.get_arguments_candidate_set_event_handler(*self._out_state_arguments[h]))
            events.append(ev)
            # tag_activator.py:291-316: the running handlers of the trashed tags
            for t in self._trash[h.tagger]:
                if t in self._running:
                    self.scheduler.trash_event(self.handlers[t])
                    self._running.discard(t)
            self._preceding = h.tagger
            done = self.system.active()[4]
        return np.array(events, dtype=L.EVENT_DTYPE)
