"""The persistent solver's tile-queue protocol (csrc/tile_queue.hpp: rings
with sequence-numbered slots, claims by compare-and-swap, queued bits,
termination by counting) proven on the CPU: tests/tile_queue_sim.cpp runs
it in host threads on shortest-path relaxations with a known answer and
checks exact values, drained rings and exact termination."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tile_queue_protocol_in_host_threads(tmp_path):
    exe = str(tmp_path / "tile_queue_sim")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-Wall", "-o", exe,
                    os.path.join(HERE, "tile_queue_sim.cpp")], check=True)
    threads = str(min(8, os.cpu_count() or 1))
    out = subprocess.run([exe, threads], capture_output=True, text=True, timeout=600)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().splitlines()[-1] == "ok"
