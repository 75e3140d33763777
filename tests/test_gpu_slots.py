"""pcbgpu_configure's slots and state limit (src/gpu.cu:377-391, 550-561; src/router.cpp:2166-2191, 2709), each in a
fresh process (tests/slots_check.py): searches from several threads at once equal their one-thread results; while
every slot is held, a solve with wait = 0 returns busy (-3) and one with wait = 1 waits and gets the same result;
a rectangle of more states than max_states is refused (-1) with the reason, and one at the limit is solved."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("slots", [1, 2])
def test_slots_busy_and_state_limit(slots):
    print("slots check", slots, flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "slots_check.py"), str(slots)], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(out, flush=True)
    assert out["concurrent_equal"]
    assert out["busy_seen"] >= 1 and out["waited_equal"]
    assert out["over_limit_rc"] == -1 and "state limit" in out["over_limit_error"]
    assert out["at_limit_equal"]
