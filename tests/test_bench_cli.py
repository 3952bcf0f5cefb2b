"""bench.py's command line without a GPU: more than one GPU is refused with a message (this build
measures one GPU and shards by shot), and every mode is listed."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr


def test_modes_are_listed():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--help'], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for flag in ('--shots', '--schedule', '--pipelines', '--density', '--gradient-2d', '--forward-2d', '--workers', '--pml', '--ot4',
                 '--copy-check', '--no-cpu', '--power2', '--autograd', '--learned', '--checkpoint', '--density-pair'):
        assert flag in r.stdout, flag

