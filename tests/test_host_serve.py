"""Host-side rules of the request server that need no device."""
import numpy as np
import pytest

from qvnt_amd import op
from qvnt_amd.serve import Server


def test_duplicate_live_id_is_refused_before_the_library_is_called():
    # a server object with one request in flight under id 7 (no device
    # handle: the refusal must come before any library call)
    s = Server.__new__(Server)
    s._h = None
    first_outs = np.zeros(4, np.uint64)
    s._live = {7: (np.zeros(4), first_outs, op.x(1))}
    with pytest.raises(ValueError, match="already in flight"):
        s.submit(7, 3, op.x(1), 1, np.zeros(4))
    # the first request's buffers are untouched and still owned
    assert s._live[7][1] is first_outs
