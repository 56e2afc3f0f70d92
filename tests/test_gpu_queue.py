"""BrokerWorkDistributor (include/mustard_queue.h, the reference's
mustard/broker_queue.h:12-150) on the GPU: ids enqueued by producer
workgroups come out exactly once through dequeue, with a ring that holds
them all and with rings of 256 and 7 slots reused while producers and
consumers race (tests/native/broker_queue_check.hip)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_broker_queue_delivers_every_id_once():
    exe = os.path.join(ROOT, "mustard_amd", "bin", "broker_queue_check")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    lines = p.stdout.splitlines()
    assert p.returncode == 0, p.stdout + p.stderr
    assert [l.split(":")[0] for l in lines] == ["fill", "mixed", "mixed_tiny_ring"]
    assert all(l.endswith("-> ok") for l in lines)
