"""Queue lifetime in the Python API: objects created on a queue keep it alive past close(), and
the garbage collector may finalise a queue and its objects in any order."""
import gc

import pytest

from bms2_amd import api

pytestmark = pytest.mark.gpu


def test_closed_queue_outlives_its_objects():
    q = api.Queue(0)
    buf = api.DeviceBuffer(q, 64)
    cand = api.Candidates(q, 4, 64)
    cycle = [buf, cand]
    cycle.append(cycle)  # only the collector can free them
    del buf, cand
    q.close()
    assert q.h  # two objects still depend on it
    del cycle
    gc.collect()
    assert q.h is None  # destroyed by the last of them
    q.close()  # idempotent


def test_queue_and_objects_in_one_garbage_cycle():
    q = api.Queue(0)
    objs = [api.DeviceBuffer(q, 128), api.Candidates(q, 8, 128), api.GMCR(q), q]
    objs.append(objs)
    q.close()
    del q, objs
    gc.collect()  # finalisers run in the collector's order; none may use a destroyed queue
    # the library still works afterwards
    q2 = api.Queue(0)
    b = api.DeviceBuffer(q2, 8)
    b.free()
    q2.close()
    assert q2.h is None
