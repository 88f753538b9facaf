"""The C ABI from a plain C caller (tests/c_abi/exec_bufs_example.c): the
mpk2 prep, the rank%d.bufs round trip and, on a GPU, mpkx_exec_bufs for all
ranks, checked in C against k x SpMV."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "tests", "c_abi", "exec_bufs_example")


def _exe():
    if not os.path.exists(EXE):
        pytest.skip("C example not built (run __graft_entry__.build())")
    return EXE


def test_c_caller_prep(tmp_path):
    r = subprocess.run([_exe(), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "prep ok" in r.stdout
    assert sorted(os.listdir(tmp_path)) == [f"rank{i}.bufs" for i in range(8)]


@pytest.mark.gpu
def test_c_caller_exec_on_gpu(tmp_path):
    r = subprocess.run([_exe(), str(tmp_path), "--gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "exec ok" in r.stdout
